"""The serial Earth driver's CSV output
(cubedsphere/tests/test_cubedsphere_dipole_background.c:120-193): the
six-face snapshot and the output frame with its index line.

The reference's own writers are compiled from the driver's source into the
harness library (oracle/ref_snapshot.c). The oracle's restatement
(orc_write_snapshot) must write the same bytes on the same state (CPU), and
libgcs.so's device-computed snapshot must write the same bytes as the
reference's writer on the GPU problem's state copied into a reference
problem (GPU)."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import (REF_LIB, ROOT, active_view, bits_equal, gcs_config, oracle_from,
                      problem, reference_from)

CP = ctypes.c_char_p
VP = ctypes.c_void_p


def declare_ref(ref):
    ref.refh_write_snapshot.argtypes = [VP, CP]
    ref.refh_write_drive_frame.argtypes = [VP, CP, ctypes.c_long, ctypes.c_long, ctypes.c_double,
                                           CP]


def read(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.fixture(scope="module")
def ref_lib():
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    lib = ctypes.CDLL(REF_LIB)
    declare_ref(lib)
    return lib


def test_oracle_snapshot_is_the_drivers_bytes(ref_lib, orc, tmp_path):
    cfg = gcs_config.workload("small")
    r = problem.Problem(ref_lib, "refh", cfg)
    r.advance(3)
    o = oracle_from(r, orc, cfg)
    orc.orc_write_snapshot.argtypes = [VP, CP]
    a, b = tmp_path / "ref.csv", tmp_path / "orc.csv"
    assert ref_lib.refh_write_snapshot(r._handle, str(a).encode()) == 0
    assert orc.orc_write_snapshot(o._handle, str(b).encode()) == 0
    text = read(a)
    lines = text.splitlines()
    assert lines[0] == b"face,i,j,k,x,y,z,rho,p,vx,vy,vz,bx,by,bz"
    assert len(lines) == 1 + cfg.active_cells
    assert lines[1].startswith(b"0,0,0,0,") and lines[-1].startswith(
        f"5,{cfg.radial_cells - 1},{cfg.angular_cells - 1},{cfg.angular_cells - 1},".encode())
    assert read(b) == text


def test_drive_frames_and_index(ref_lib, tmp_path):
    """Two frames and their index lines (:177-193, header from :561)."""
    cfg = gcs_config.workload("tiny")
    r = problem.Problem(ref_lib, "refh", cfg)
    prefix, index = tmp_path / "drive", tmp_path / "drive.index.csv"
    for frame, step, seconds in ((0, 1, 0.5), (1, 7, 600.25)):
        r.advance(1)
        assert ref_lib.refh_write_drive_frame(r._handle, str(prefix).encode(), frame, step,
                                              seconds, str(index).encode()) == 0
    lines = read(index).decode().splitlines()
    assert lines == ["frame,step,simulation_seconds,mhd_csv",
                     f"0,1,0.5,{prefix}_000000.csv", f"1,7,600.25,{prefix}_000001.csv"]


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps", [("small", 6), ("ragged", 4)])
def test_device_snapshot_is_the_drivers_bytes(gpu, name, steps, tmp_path):
    """gcs_problem_write_snapshot (primitives and total B computed on the
    device, formatted on the host) against the driver's writer on the same
    state, byte for byte; and the drive frame's CSV and index line."""
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    ref = ctypes.CDLL(REF_LIB)
    declare_ref(ref)
    cfg = gcs_config.workload(name)
    g = gpu.GpuProblem(cfg)
    g.advance(steps)
    lib = gpu.library()
    lib.gcs_problem_write_snapshot.argtypes = [VP, CP]
    lib.gcs_problem_write_drive_frame.argtypes = [VP, CP, ctypes.c_long, ctypes.c_long,
                                                  ctypes.c_double, CP]
    ours, theirs = tmp_path / "gcs.csv", tmp_path / "ref.csv"
    assert lib.gcs_problem_write_snapshot(g._handle, str(ours).encode()) == 0, gpu.last_error()
    r = reference_from(g, ref, cfg)
    assert ref.refh_write_snapshot(r._handle, str(theirs).encode()) == 0
    a, b = read(ours), read(theirs)
    lines = a.count(b"\n")
    print(f"snapshot {name}: {len(a)} bytes, {lines} lines", flush=True)
    assert a == b
    # an output frame: the same CSV under the frame name and its index line
    prefix, index = tmp_path / "frame", tmp_path / "frame.index.csv"
    assert lib.gcs_problem_write_drive_frame(g._handle, str(prefix).encode(), 3, steps, 1.25,
                                             str(index).encode()) == 0
    assert read(f"{prefix}_000003.csv") == b
    assert read(index).decode().splitlines() == [
        "frame,step,simulation_seconds,mhd_csv", f"3,{steps},1.25,{prefix}_000003.csv"]


def load_ids():
    return ((problem.CONSERVED, 5), (problem.OLD_CONSERVED, 5), (problem.CELL_MAGNETIC, 3),
            (problem.OLD_CELL_MAGNETIC, 3))


def mi_csv_files(prefix):
    return {ext: open(f"{prefix}.{ext}").read().splitlines()
            for ext in ("mi_north.csv", "mi_south.csv", "shell_j.csv")}


def assert_mi_csv_close(want, got, rel):
    """The ionosphere's CSV files: identical headers, grid columns and row
    counts; values within `rel` of each column's largest magnitude."""
    for ext in want:
        exact_columns = 3 if ext == "shell_j.csv" else 4
        fa, fb = want[ext], got[ext]
        assert fa[0] == fb[0] and len(fa) == len(fb), ext
        ra = [line.split(",") for line in fa[1:]]
        rb = [line.split(",") for line in fb[1:]]
        assert all(x[:exact_columns] == y[:exact_columns] for x, y in zip(ra, rb)), ext
        va = np.array([[float(v) for v in x[exact_columns:]] for x in ra])
        vb = np.array([[float(v) for v in x[exact_columns:]] for x in rb])
        scale = np.maximum(np.abs(va).max(axis=0), 1e-300)
        err = (np.abs(va - vb).max(axis=0) / scale).max()
        print(f"{ext}: {len(ra)} rows, largest column-relative difference {err:.3e}", flush=True)
        assert err <= rel, ext


@pytest.mark.gpu
def test_device_load_is_the_drivers_load(gpu, tmp_path):
    """gcs_problem_load_snapshot against the driver's load_mhd_snapshot
    (:61-118) on the same file: the active conserved state and residual
    cell B of both levels bit for bit; then the earth-mi-replay sequence
    (:526-533), one ionosphere update at t = 0 and its CSV files, against
    the reference's ionosphere on the reference's loaded state."""
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    ref = ctypes.CDLL(REF_LIB)
    declare_ref(ref)
    ref.refh_load_snapshot.argtypes = [VP, CP]
    ref.refh_mi.restype = VP
    ref.refh_mi.argtypes = [VP]
    ref.gamera_cs_mi_dipole_maybe_update.argtypes = [VP, ctypes.c_double]
    ref.gamera_cs_mi_dipole_write_csv.argtypes = [VP, CP]
    ref.gamera_cs_mi_dipole_write_shell_current_csv.argtypes = [VP, CP]
    lib = gpu.library()
    lib.gcs_problem_write_snapshot.argtypes = [VP, CP]
    lib.gcs_problem_load_snapshot.argtypes = [VP, CP]
    lib.gcs_problem_mi_maybe_update.argtypes = [VP, ctypes.c_double,
                                                ctypes.POINTER(ctypes.c_int)]
    lib.gcs_problem_write_mi_csv.argtypes = [VP, CP]
    cfg = gcs_config.workload("small_mi")
    cfg.mi.relative_tolerance = 1e-13
    src = gpu.GpuProblem(cfg)
    src.advance(8)
    path = tmp_path / "state.csv"
    assert lib.gcs_problem_write_snapshot(src._handle, str(path).encode()) == 0, gpu.last_error()
    g = gpu.GpuProblem(cfg)
    r = problem.Problem(ref, "refh", cfg)
    print("load small_mi snapshot of step 8", flush=True)
    assert lib.gcs_problem_load_snapshot(g._handle, str(path).encode()) == 0, gpu.last_error()
    assert ref.refh_load_snapshot(r._handle, str(path).encode()) == 0
    for f in range(gcs_config.FACE_COUNT):
        for fid, nc in load_ids():
            a, b = active_view(g, f, fid, cfg, nc), active_view(r, f, fid, cfg, nc)
            assert bits_equal(a, b), (f, fid)
    # the loaded state is the written one up to the primitive round trip
    rho_src = active_view(src, 0, problem.CONSERVED, cfg, 5)[..., 0]
    rho_new = active_view(g, 0, problem.CONSERVED, cfg, 5)[..., 0]
    assert np.abs(rho_src - rho_new).max() <= 1e-14 * np.abs(rho_src).max()
    updated = ctypes.c_int(-1)
    assert lib.gcs_problem_mi_maybe_update(g._handle, 0.0, ctypes.byref(updated)) == 0
    assert updated.value == 1
    ref_mi = ref.refh_mi(r._handle)
    assert ref.gamera_cs_mi_dipole_maybe_update(ref_mi, 0.0) == 1
    ours, theirs = tmp_path / "gcs", tmp_path / "ref"
    assert lib.gcs_problem_write_mi_csv(g._handle, str(ours).encode()) == 0, gpu.last_error()
    assert ref.gamera_cs_mi_dipole_write_csv(ref_mi, str(theirs).encode()) == 0
    assert ref.gamera_cs_mi_dipole_write_shell_current_csv(ref_mi, str(theirs).encode()) == 0
    # the shell current is sampled from identical B; the potential is the
    # Krylov solution at relative tolerance 1e-13
    assert_mi_csv_close(mi_csv_files(theirs), mi_csv_files(ours), 1e-9)


@pytest.mark.gpu
def test_device_load_rejects_what_the_driver_rejects(gpu, tmp_path):
    """Files the driver's loader rejects (wrong header, a missing line, a
    cell index out of range, a centroid moved by 1e-6, a density the
    conversion refuses, a line longer than its read buffer) make
    gcs_problem_load_snapshot fail with the state unchanged; the
    reference's loader returns -1 on each as well."""
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    ref = ctypes.CDLL(REF_LIB)
    ref.refh_load_snapshot.argtypes = [VP, CP]
    lib = gpu.library()
    lib.gcs_problem_write_snapshot.argtypes = [VP, CP]
    lib.gcs_problem_load_snapshot.argtypes = [VP, CP]
    cfg = gcs_config.workload("tiny")
    g = gpu.GpuProblem(cfg)
    g.advance(2)
    good = tmp_path / "good.csv"
    assert lib.gcs_problem_write_snapshot(g._handle, str(good).encode()) == 0, gpu.last_error()
    lines = read(good).decode().splitlines(keepends=True)

    def edit(row, column, value):
        out = list(lines)
        cells = out[row].rstrip("\n").split(",")
        cells[column] = value(cells[column])
        out[row] = ",".join(cells) + "\n"
        return out

    cases = {
        "header": ["face,i,j,k,x,y,z,rho,p,vx,vy,vz,bx,by\n"] + lines[1:],
        "missing line": lines[:-1],
        "face 6": edit(5, 0, lambda v: "6"),
        "i out of range": edit(7, 1, lambda v: str(cfg.radial_cells)),
        "centroid": edit(len(lines) // 2, 5, lambda v: repr(float(v) + 1e-6)),
        "density": edit(len(lines) - 3, 7, lambda v: "nan"),
        # longer than the driver's 1024-byte fgets buffer: the reference
        # parses the line's head and fails on its remainder
        "long line": lines[:9] + [lines[9].rstrip("\n") + " " * 1100 + "\n"] + lines[10:],
    }
    before = [g.field(f, problem.CONSERVED).copy() for f in range(gcs_config.FACE_COUNT)]
    r = problem.Problem(ref, "refh", cfg)
    for name, text in cases.items():
        path = tmp_path / "bad.csv"
        path.write_text("".join(text))
        print(f"reject {name}", flush=True)
        assert lib.gcs_problem_load_snapshot(g._handle, str(path).encode()) == -1, name
        assert "snapshot load" in gpu.last_error(), name
        assert ref.refh_load_snapshot(r._handle, str(path).encode()) == -1, name
        for f in range(gcs_config.FACE_COUNT):
            assert bits_equal(g.field(f, problem.CONSERVED), before[f]), (name, f)
    assert lib.gcs_problem_load_snapshot(g._handle, str(good).encode()) == 0, gpu.last_error()


def perturbed_snapshot(src, dst, amplitude=20.0):
    """A copy of the snapshot `src` whose total cell B carries a smooth
    twisted perturbation (amplitude e^{-(r - 2.5)/2} / r (-y, x (1 + 0.3 z/r), 0)),
    so that the field-aligned current on the MI shell is well above
    roundoff; centroids, indices and fluid state are unchanged."""
    with open(src) as f:
        lines = f.read().splitlines()
    out = [lines[0]]
    for line in lines[1:]:
        c = line.split(",")
        x, y, z = (float(v) for v in c[4:7])
        r = math.sqrt(x * x + y * y + z * z)
        scale = amplitude * math.exp(-(r - 2.5) / 2.0) / r
        b = [float(v) for v in c[12:15]]
        b[0] += -y * scale
        b[1] += x * scale * (1.0 + 0.3 * z / r)
        c[12:15] = [repr(v) for v in b]
        out.append(",".join(c))
    with open(dst, "w") as f:
        f.write("\n".join(out) + "\n")


@pytest.mark.gpu
def test_driver_mi_replay_through_dropin(tmp_path):
    """The Earth driver's earth-mi-replay mode (:526-533) built with
    libgcs.so's ionosphere (oracle/Makefile `programs`) against the
    reference build: the driver's own loader fills the host storage, and
    the drop-in's maybe_update uploads it, samples the shell current and
    solves on the GPU. The snapshot is the reference build's state after one
    step (72 x 10 x 10 per face) with a twisted B perturbation. The driver's
    ionosphere solves to relative residual 1e-8 (:512-518), so potentials
    agree to 1e-6 of their largest magnitude; the sampled currents, from the
    same B, to 1e-10."""
    bin_dir = os.path.join(ROOT, "oracle", "_ref", "bin")
    exe = os.path.join(bin_dir, "test_cubedsphere_dipole_background")
    if not (os.path.exists(exe) and os.path.exists(exe + "_gcs")):
        pytest.skip("reference programs not built (make -C oracle programs)")
    env = dict(os.environ, GAMERA_CS_EARTH_ANGULAR_CELLS="10")
    print("driver earth-mi-local 1 (reference build)", flush=True)
    run = subprocess.run([exe, "earth-mi-local", "1", "state.csv"], cwd=tmp_path, env=env,
                         capture_output=True, text=True, timeout=240)
    assert run.returncode == 0, run.stderr[-1000:]
    perturbed_snapshot(tmp_path / "state.csv", tmp_path / "twisted.csv")
    for program, prefix in ((exe, "ref"), (exe + "_gcs", "gcs")):
        print(f"driver earth-mi-replay twisted.csv {prefix}", flush=True)
        run = subprocess.run([program, "earth-mi-replay", "twisted.csv", prefix], cwd=tmp_path,
                             env=env, capture_output=True, text=True, timeout=240)
        assert run.returncode == 0, (prefix, run.stdout[-1000:], run.stderr[-1000:])
    want, got = mi_csv_files(tmp_path / "ref"), mi_csv_files(tmp_path / "gcs")
    assert_mi_csv_close({"shell_j.csv": want["shell_j.csv"]}, {"shell_j.csv": got["shell_j.csv"]},
                        1e-10)
    assert_mi_csv_close(want, got, 1e-6)
