"""Record the output of the reference's own step tests and Earth driver.

Runs the programs built by `make -C oracle programs` with the reference's
step (oracle/_ref/bin/test_cubedsphere_X, CPU) for the invocations the
reference registers as tests (cubedsphere/CMakeLists.txt:103-160) and
writes their exit codes and standard output to programs.json. The GPU test
runs the same programs built against libgcs.so (..._gcs) and compares.

    python tests/golden/make_program_golden.py
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(ROOT, "oracle", "_ref", "bin")

# (case name, program, arguments): the registered invocations whose runs
# take well under a minute on one node.
CASES = [
    ("uniform_mhd", "uniform_mhd", []),
    ("six_mhd", "six_mhd", []),
    ("mhd_blast", "mhd_blast", []),
    ("mhd_shock", "mhd_shock", []),
    ("alfven", "alfven", []),
    ("dipole_background", "dipole_background", []),
    ("dipole_background_72x16x16", "dipole_background", ["earth-local"]),
    ("earth_mi_dipole_zero_fac", "dipole_background", ["earth-mi-local", "1"]),
    # hydrodynamic steps (use_mhd = false) and the face-centred MHD blast
    # (cubedsphere/CMakeLists.txt:139-147)
    ("hd_spherical_blast", "hd_spherical_blast", []),
    ("hd_spherical_blast_six_faces", "hd_spherical_blast", ["-", "24", "1.5"]),
    ("mhd_spherical_blast_face_center", "hd_spherical_blast",
     ["-", "32", "1.5", "face-center-mhd", "0.1"]),
]


def main():
    out = {}
    for name, prog, args in CASES:
        exe = os.path.join(BIN, f"test_cubedsphere_{prog}")
        if not os.path.exists(exe):
            sys.exit(f"{exe} missing: run make -C oracle programs")
        print(f"{name}: {prog} {' '.join(args)}", flush=True)
        run = subprocess.run([exe, *args], capture_output=True, text=True,
                             timeout=600)
        out[name] = {"program": prog, "args": args, "rc": run.returncode,
                     "stdout": run.stdout}
        print(f"  rc={run.returncode} {run.stdout.strip().splitlines()[-1:]}",
              flush=True)
    with open(os.path.join(HERE, "programs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
