"""The HIP step against the CPU oracle and the compiled reference, on the
GPU. Setup must be bitwise the reference's; steps agree to the reference's
scaled parity measure |a - b| / max(1, |a|) (tolerances below)."""
import ctypes
import time

import numpy as np
import pytest

from conftest import (bits_equal, counters, gcs_config, max_divergence,
                      oracle_from, problem, scaled_error, set_wind, wind_series)

pytestmark = pytest.mark.gpu

# One step differs from the oracle only by last-ulp differences of ROCm's
# erfc/exp/acos/hypot from glibc's; 1e-13 leaves two decades of margin.
STEP_TOL = 1e-13
# After the configured 200 steps (config 0): the contract's 1e-12.
RUN_TOL = 1e-12

STATE = ([problem.CONSERVED, problem.OLD_CONSERVED, problem.CELL_MAGNETIC,
          problem.OLD_CELL_MAGNETIC]
         + [problem.FACE_FLUX + d for d in range(3)]
         + [problem.OLD_FACE_FLUX + d for d in range(3)])


def worst_error(a, b, fields):
    worst = 0.0
    for fid in fields:
        for f in range(gcs_config.FACE_COUNT):
            worst = max(worst, scaled_error(a.field(f, fid), b.field(f, fid)))
    return worst


@pytest.mark.parametrize("name", ["tiny", "small", "ragged", "ghost5"])
def test_setup_is_bitwise_the_reference_setup(gpu, ref, name):
    cfg = gcs_config.workload(name)
    print(f"setup parity {name}", flush=True)
    g = gpu.GpuProblem(cfg)
    r = problem.Problem(ref, "refh", cfg)
    for fid in problem.GEOMETRY_FIELDS + problem.BACKGROUND_FIELDS + problem.STATE_FIELDS:
        for f in range(gcs_config.FACE_COUNT):
            assert bits_equal(r.field(f, fid), g.field(f, fid)), (fid, f)
    for f in range(gcs_config.FACE_COUNT):
        for d in range(3):
            assert np.array_equal(r.edge_valid(f, d), g.edge_valid(f, d))


@pytest.mark.parametrize("name,steps", [("tiny", 10), ("small", 10), ("ragged", 10),
                                        ("ghost5", 6)])
def test_steps_match_oracle(gpu, orc, name, steps):
    cfg = gcs_config.workload(name)
    print(f"step parity {name} x{steps}", flush=True)
    g = gpu.GpuProblem(cfg)
    o = oracle_from(g, orc, cfg)
    checked = STATE + [problem.EDGE_EMF + d for d in range(3)] + [
        problem.PREDICTED_CONSERVED, problem.PREDICTED_CELL_MAGNETIC]
    for s in range(steps):
        g.advance(1)
        o.advance(1)
        cg, co = g.clock(), o.clock()
        assert cg["steps"] == co["steps"]
        assert abs(cg["dt"] - co["dt"]) <= 1e-14 * co["dt"]
        err = worst_error(o, g, checked)
        assert err <= STEP_TOL * (s + 1), (s, err)
    assert max_divergence(g, cfg) < 1e-10


def test_emergency_diffusion_steps_match_oracle(gpu, orc):
    """A 16 code-speed wind against the Boris light speed 10: interface
    speeds reach 1.5 c and emergency diffusion (nonorthogonal_flux.c:206-241)
    fires on most faces. State, predicted state, EMF and the per-face
    counters (nonorthogonal_sweep.c:541-545) against the oracle every step."""
    cfg = gcs_config.workload("fast_wind")
    print("step parity fast_wind x6 (emergency diffusion)", flush=True)
    g = gpu.GpuProblem(cfg)
    o = oracle_from(g, orc, cfg)
    checked = STATE + [problem.EDGE_EMF + d for d in range(3)] + [
        problem.PREDICTED_CONSERVED, problem.PREDICTED_CELL_MAGNETIC]
    for s in range(6):
        g.advance(1)
        o.advance(1)
        assert abs(g.clock()["dt"] - o.clock()["dt"]) <= 1e-14 * o.clock()["dt"]
        err = worst_error(o, g, checked)
        assert err <= STEP_TOL * (s + 1), (s, err)
        for f in range(gcs_config.FACE_COUNT):
            (cg, vg), (co, vo) = counters(g, f), counters(o, f)
            assert cg == co, (s, f, cg, co)
            assert abs(vg - vo) <= 1e-13 * vo, (s, f, vg, vo)
    fired = sum(counters(g, f)[0] for f in range(gcs_config.FACE_COUNT))
    print(f"fast_wind: emergency diffusion on {fired:.0f} faces, scaled error {err:.2e}",
          flush=True)
    assert fired > 0
    assert max_divergence(g, cfg) < 1e-10


@pytest.mark.parametrize("linear", [True, False])
def test_long_wind_series_steps_match_oracle(gpu, orc, linear):
    """A 40-sample solar-wind series (any length, solar_wind.c:55-225) set on
    the device problem and on the oracle: ballistic sampling, wind weights,
    the outer ghosts and EMF over 8 steps."""
    cfg = gcs_config.workload("small")
    print(f"long wind series, linear={linear}", flush=True)
    g = gpu.GpuProblem(cfg)
    o = oracle_from(g, orc, cfg)
    time, state = wind_series()
    set_wind(g, time, state, linear)
    set_wind(o, time, state, linear)
    for s in range(8):
        g.advance(1)
        o.advance(1)
        err = worst_error(o, g, STATE)
        assert err <= STEP_TOL * (s + 1), (s, err)
    print(f"  scaled error after 8 steps {err:.2e}", flush=True)


def test_explicit_step_matches_oracle(gpu, orc):
    cfg = gcs_config.workload("small")
    g = gpu.GpuProblem(cfg)
    o = oracle_from(g, orc, cfg)
    for ratio, dt in ((0.0, 2e-3), (0.37, 3e-3)):
        g.step_explicit(ratio, dt, 0.1, 0.097)
        o.step_explicit(ratio, dt, 0.1, 0.097)
        assert worst_error(o, g, STATE) <= STEP_TOL


def test_gpu_step_is_deterministic(gpu):
    cfg = gcs_config.workload("small")
    a, b = gpu.GpuProblem(cfg), gpu.GpuProblem(cfg)
    a.advance(5)
    b.advance(5)
    for fid in STATE:
        for f in range(gcs_config.FACE_COUNT):
            assert bits_equal(a.field(f, fid), b.field(f, fid))


def test_config0_200_steps_match_oracle(gpu, orc):
    """Config 0: 6 x 24^2 x 48, 200 steps of Up7/PDM + flux + CT."""
    cfg = gcs_config.workload("config0")
    g = gpu.GpuProblem(cfg)
    o = oracle_from(g, orc, cfg)
    t0 = time.time()
    for chunk in range(4):
        g.advance(50)
        o.advance(50)
        err = worst_error(o, g, STATE[:4])
        print(f"config0 step {50 * (chunk + 1)} scaled error {err:.3e} "
              f"({time.time() - t0:.0f}s)", flush=True)
        assert err <= RUN_TOL
    assert abs(g.clock()["time"] - o.clock()["time"]) <= 1e-13 * o.clock()["time"]
    assert max_divergence(g, cfg) < 1e-10
    for f in range(gcs_config.FACE_COUNT):
        u = g.field(f, problem.CONSERVED).reshape(-1, 5)
        assert np.all(u[:, 0] > 0)


def test_final_gates_match_the_drivers_gate(gpu, orc):
    """gcs_problem_final_gates against the Earth driver's end-of-run gate
    (test_cubedsphere_dipole_background.c:807-893) evaluated on the oracle's
    state after the same 20 steps: minimum density and pressure of the
    primitive conversion with floors 1e-12, maximum total div B."""
    cfg = gcs_config.workload("small")
    g = gpu.GpuProblem(cfg)
    o = oracle_from(g, orc, cfg)
    g.advance(20)
    o.advance(20)
    out = (ctypes.c_double * 3)()
    fn = gpu.library().gcs_problem_final_gates
    fn.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    assert fn(g._handle, out) == 0
    ni, nj, nk = cfg.extent
    gh, nr, n = cfg.ghost, cfg.radial_cells, cfg.angular_cells
    gamma = cfg.startup.gamma
    rho_min = p_min = np.inf
    for f in range(gcs_config.FACE_COUNT):
        u = o.field(f, problem.CONSERVED).reshape(ni, nj, nk, 5)[
            gh:gh + nr, gh:gh + n, gh:gh + n].reshape(-1, 5)
        v = u[:, 1:4] / u[:, 0:1]
        v2 = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
        p = np.maximum((gamma - 1.0) * (u[:, 4] - 0.5 * u[:, 0] * v2), 1e-12)
        rho_min = min(rho_min, float(np.maximum(u[:, 0], 1e-12).min()))
        p_min = min(p_min, float(p.min()))
    div = max_divergence(o, cfg)
    print(f"gates: GPU {list(out)}; oracle state ({rho_min}, {p_min}, {div})", flush=True)
    assert abs(out[0] - rho_min) <= 1e-12 * rho_min
    assert abs(out[1] - p_min) <= 1e-12 * p_min
    assert abs(out[2] - div) <= 1e-6 * div
    assert out[0] > 0 and out[1] > 0 and out[2] < 1e-9
