"""Throughput of the MI355X six-face GAMERA-Cube step.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--gpus 1]

Workload (default `config1_nomi`): 6 faces x 96^2 angular x 192 radial cells,
log radial grid 2.5-30 R_E, uniform solar wind with 1% density/pressure
perturbations, aligned dipole, Earth radial boundary with the ionosphere
potential held at zero. The state is resident in HBM before timing starts.

Prints one JSON line: cell-updates/s over K driver steps (CFL dt rule + full
six-face step), the step's fraction of the HBM roofline under the
1216 B/cell-update model, the dominant kernel's roofline from HIP events on
the step's stream, and the compiled reference library's CPU rate on the
reference's own CPU-runnable case (6 x 24^2 x 48).
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Compulsory bytes of one cell-update of the whole step (FP64; 30 doubles of
# state + 122 of static geometry/background).
STEP_BYTES_PER_CELL = 1216
# Algorithmic (compulsory) bytes per unit of each heavy stage, and the unit
# count per launch as a function of (nr, n); see DESIGN.md "Kernels".
STAGE_MODEL = {
    "face_flux_i": (256, lambda nr, n: 6 * (nr + 1) * n * n),
    "face_flux_j": (256, lambda nr, n: 6 * nr * (n + 1) * n),
    "face_flux_k": (256, lambda nr, n: 6 * nr * n * (n + 1)),
    "edge_emf_i": (184, lambda nr, n: 6 * nr * (n + 1) * (n + 1)),
    "edge_emf_j": (184, lambda nr, n: 6 * (nr + 1) * n * (n + 1)),
    "edge_emf_k": (184, lambda nr, n: 6 * (nr + 1) * (n + 1) * n),
    "update": (616, lambda nr, n: 6 * nr * n * n),
    "predict": (392, lambda nr, n: 6 * (nr + 9) * (n + 9) * (n + 9)),
}


def cpu_baseline(steps=20):
    """Compiled reference library on the reference's CPU case (config 0)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libgamera_ref.so")
    if not os.path.exists(path):
        return None
    cfgm = importlib.import_module("gamera-cube_amd.config")
    prob = importlib.import_module("gamera-cube_amd.problem")
    lib = ctypes.CDLL(path)
    cfg = cfgm.workload("config0")
    p = prob.Problem(lib, "refh", cfg)
    p.advance(1)
    t0 = time.perf_counter()
    p.advance(steps)
    seconds = time.perf_counter() - t0
    p.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": cfg.active_cells * steps / seconds, "unit": "cell-updates/s",
            "cores": cores, "kind": "reference",
            "sample": f"6x24^2x48 (config 0), {steps} steps of the compiled "
                      f"reference serial library (OpenMP), {seconds:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="config1_nomi")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs one GPU: the six-face step is a single-GPU "
              "build; --gpus above 1 is not supported", file=sys.stderr)
        return 2
    cfgm = importlib.import_module("gamera-cube_amd.config")
    native = importlib.import_module("gamera-cube_amd.native")
    cfg = cfgm.workload(args.workload)
    print(f"bench: workload {args.workload} {cfg.active} x 6, setup ...",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    g = native.GpuProblem(cfg)
    g.set_fidelity(False)
    print(f"bench: setup {time.perf_counter() - t0:.1f} s, warmup {args.warmup}",
          file=sys.stderr, flush=True)
    if args.warmup:
        g.advance(args.warmup)
    ms, stages = g.time_steps(args.steps, profile=True)
    print(f"bench: {args.steps} steps in {ms:.1f} ms", file=sys.stderr, flush=True)
    per_step = ms / args.steps
    value = cfg.active_cells * args.steps / (ms * 1e-3)
    nr, n = cfg.radial_cells, cfg.angular_cells
    modeled = {k: v for k, v in stages.items() if k in STAGE_MODEL and v > 0}
    top = max(modeled, key=modeled.get)
    bytes_per_unit, units = STAGE_MODEL[top]
    launch_ms = stages[top] / args.steps
    achieved = bytes_per_unit * units(nr, n) / (launch_ms * 1e-3) / 1e9
    result = {
        "metric": "cell-updates/s on 1 MI355X at 6x96^2x192",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload,
                   "cells": f"6x{n}^2x{nr}",
                   "active_cells": cfg.active_cells,
                   "mi": "ionosphere potential held at zero (MI solve not in this build)",
                   "hbm_fraction_step_model": STEP_BYTES_PER_CELL * value / (HBM_PEAK_GBS * 1e9)},
        "roofline": {"bound": "hbm", "kernel": top,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "launch_ms": launch_ms, "bytes_per_unit": bytes_per_unit,
                     "units_per_launch": units(nr, n)},
        "stage_ms_per_step": {k: round(v / args.steps, 4) for k, v in stages.items() if v > 0},
        "cpu_baseline": None,
    }
    g.close()
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (compiled reference, config 0) ...",
              file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
