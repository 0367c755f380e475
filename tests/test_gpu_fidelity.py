"""Fidelity 0 (gcs_problem_set_fidelity, what bench.py runs) against the
default fidelity 1 (the reference's storage, ghost values included).

Fidelity 0 skips work that only reproduces the reference's ghost values in
storage: the new level's ghost copy and recovery after the update, the
ghost-cell B of the halo fill and the ghost-cell B recovery of the Earth
boundary. Nothing the step computes reads those values, so the active
state, the clock and the ionosphere potential must be equal bit for bit.
"""
import numpy as np
import pytest

from conftest import bits_equal, gcs_config, problem

pytestmark = pytest.mark.gpu


def active_cells(p, cfg, fid, f):
    ni, nj, nk = cfg.extent
    g, nr, n = cfg.ghost, cfg.radial_cells, cfg.angular_cells
    ncomp = 5 if fid in (problem.CONSERVED, problem.OLD_CONSERVED) else 3
    a = p.field(f, fid).reshape(ni, nj, nk, ncomp)
    return a[g:g + nr, g:g + n, g:g + n]


def active_faces(p, cfg, d, f):
    ni, nj, nk = cfg.extent
    g, nr, n = cfg.ghost, cfg.radial_cells, cfg.angular_cells
    ext = [ni + (d == 0), nj + (d == 1), nk + (d == 2)]
    a = p.field(f, problem.FACE_FLUX + d).reshape(ext)
    return a[g:g + nr + (d == 0), g:g + n + (d == 1), g:g + n + (d == 2)]


@pytest.mark.parametrize("name,steps", [("small", 12), ("small_mi", 12)])
def test_fidelity_zero_keeps_active_results(gpu, name, steps):
    cfg = gcs_config.workload(name)
    print(f"fidelity 0 vs 1 {name} x{steps}", flush=True)
    a = gpu.GpuProblem(cfg)
    b = gpu.GpuProblem(cfg)
    b.set_fidelity(False)
    b.set_graph(True)
    for s in range(steps):
        a.advance(1)
        b.advance(1)
        assert a.clock() == b.clock(), s
        for f in range(gcs_config.FACE_COUNT):
            for fid in (problem.CONSERVED, problem.CELL_MAGNETIC):
                assert bits_equal(active_cells(a, cfg, fid, f),
                                  active_cells(b, cfg, fid, f)), (s, fid, f)
            for d in range(3):
                assert bits_equal(active_faces(a, cfg, d, f), active_faces(b, cfg, d, f)), (s, d, f)
    if cfg.mi_enabled:
        ra, rb = a.mi_report(), b.mi_report()
        assert ra["updates"] == rb["updates"] >= 1
        assert np.array_equal(a.mi_field(0), b.mi_field(0))
