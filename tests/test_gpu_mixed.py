"""The mixed-precision step (FP32 reconstruction, fluxes and EMF; FP64
update, CT and CFL) against the FP64 step on the same problem. The FP64
step is the oracle's to 1e-12 (test_gpu_parity.py); the mixed form must
stay within a stated tolerance of it and keep div B at FP64 roundoff."""
import numpy as np
import pytest

from conftest import gcs_config, max_divergence, oracle_from, problem

pytestmark = pytest.mark.gpu

# Relative L2 error of rho, p, |v| and B after the steps below (FP32 has
# ~6e-8 unit roundoff; the differences grow slowly with the steps).
MIXED_TOL = 1e-4


def primitives(p, cfg):
    ni, nj, nk = cfg.extent
    g = cfg.ghost
    out = {"rho": [], "p": [], "v": [], "b": []}
    for f in range(gcs_config.FACE_COUNT):
        u = p.field(f, problem.CONSERVED).reshape(ni, nj, nk, 5)[
            g:g + cfg.radial_cells, g:g + cfg.angular_cells, g:g + cfg.angular_cells]
        b = p.field(f, problem.CELL_MAGNETIC).reshape(ni, nj, nk, 3)[
            g:g + cfg.radial_cells, g:g + cfg.angular_cells, g:g + cfg.angular_cells]
        rho = u[..., 0]
        v = u[..., 1:4] / rho[..., None]
        pr = (5.0 / 3.0 - 1.0) * (u[..., 4] - 0.5 * rho * (v * v).sum(-1))
        out["rho"].append(rho.ravel())
        out["p"].append(pr.ravel())
        out["v"].append(np.linalg.norm(v, axis=-1).ravel())
        out["b"].append(np.linalg.norm(b, axis=-1).ravel())
    return {k: np.concatenate(v) for k, v in out.items()}


@pytest.mark.parametrize("name,steps", [("small", 40), ("config0", 100)])
def test_mixed_step_close_to_fp64(gpu, name, steps):
    cfg = gcs_config.workload(name)
    print(f"mixed vs fp64 {name} x{steps}", flush=True)
    a = gpu.GpuProblem(cfg)
    b = gpu.GpuProblem(cfg)
    b.set_precision(True)
    a.advance(steps)
    b.advance(steps)
    pa, pb = primitives(a, cfg), primitives(b, cfg)
    for key in ("rho", "p", "v", "b"):
        l2 = np.linalg.norm(pb[key] - pa[key]) / np.linalg.norm(pa[key])
        linf = np.abs(pb[key] - pa[key]).max() / np.abs(pa[key]).max()
        print(f"  {key}: rel L2 {l2:.2e} rel Linf {linf:.2e}", flush=True)
        assert l2 <= MIXED_TOL
    div = max_divergence(b, cfg)
    print(f"  max |div B| (mixed) {div:.2e}", flush=True)
    assert div < 1e-9


@pytest.mark.timeout(900)
def test_mixed_form_against_oracle_config0(gpu, orc):
    """Config 0 (6 x 24^2 x 48) for its 200 steps in the mixed form against
    the FP64 oracle started from the same arrays: relative L2 / Linf error
    of rho, p, |v|, |B| and max |div B|, each with its own dt rule."""
    cfg = gcs_config.workload("config0")
    print("mixed vs oracle config0 x200", flush=True)
    b = gpu.GpuProblem(cfg)
    b.set_precision(True)
    o = oracle_from(b, orc, cfg)
    for chunk in range(4):
        b.advance(50)
        o.advance(50)
        print(f"  step {50 * (chunk + 1)}", flush=True)
    po, pb = primitives(o, cfg), primitives(b, cfg)
    worst = {}
    for key in ("rho", "p", "v", "b"):
        l2 = np.linalg.norm(pb[key] - po[key]) / np.linalg.norm(po[key])
        linf = np.abs(pb[key] - po[key]).max() / np.abs(po[key]).max()
        worst[key] = (l2, linf)
        print(f"  {key}: rel L2 {l2:.2e} rel Linf {linf:.2e}", flush=True)
        assert l2 <= MIXED_TOL
        assert linf <= 10 * MIXED_TOL
    div_b, div_o = max_divergence(b, cfg), max_divergence(o, cfg)
    print(f"  max |div B| mixed {div_b:.2e}, oracle {div_o:.2e}", flush=True)
    assert div_b < 1e-9
    assert all(np.all(pb[k] > 0) for k in ("rho", "p"))


def test_mixed_downloads_widen_fp32_arrays(gpu):
    """The mixed form keeps the predicted state and the face fluxes in FP32
    arrays; downloads widen them, so they track the FP64 step's arrays to
    FP32 accuracy (and are not the stale FP64 arrays)."""
    cfg = gcs_config.workload("small")
    a = gpu.GpuProblem(cfg)
    b = gpu.GpuProblem(cfg)
    b.set_precision(True)
    a.advance(10)
    b.advance(10)
    fields = ([problem.PREDICTED_CONSERVED, problem.PREDICTED_CELL_MAGNETIC]
              + [problem.FLUID_FACE_FLUX + d for d in range(3)]
              + [problem.MAXWELL_FACE_FLUX + d for d in range(3)])
    for fid in fields:
        worst = 0.0
        for f in range(gcs_config.FACE_COUNT):
            x, y = a.field(f, fid), b.field(f, fid)
            scale = np.abs(x).max()
            assert scale > 0.0, fid
            worst = max(worst, float(np.abs(x - y).max() / scale))
        print(f"  field {fid}: max |mixed - fp64| / max |fp64| = {worst:.2e}", flush=True)
        assert worst <= 1e-4, fid
    # switching back to FP64 downloads the FP64 arrays again
    b.set_precision(False)
    b.advance(1)
    a.advance(1)
    x, y = a.field(0, problem.PREDICTED_CONSERVED), b.field(0, problem.PREDICTED_CONSERVED)
    assert float(np.abs(x - y).max() / np.abs(x).max()) <= 1e-4
