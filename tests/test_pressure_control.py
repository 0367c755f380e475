"""Post-step pressure control (gamera_no_apply_pressure_control,
NO/nonorthogonal_step.c:74-171, as the MPI Earth driver applies it after a
step, cubedsphere/tests/test_cubedsphere_mpi_earth_tile.c:459-479).

The oracle's restatement is checked bitwise against the compiled reference
function; the device kernel (gcs_problem_pressure_control) against the
oracle. Besides the driver's own parameters, which change no cell of these
small problems, the thresholds are set from the state itself so that about
half the cells take each branch.
"""
import ctypes

import numpy as np
import pytest

from conftest import bits_equal, gcs_config, oracle_from, problem, scaled_error

DRIVER = dict(density_floor=1e-6, pressure_floor=1e-8, light_speed=10.0,
              low_density_threshold=1.0e-3, low_density_sound_speed=1.0e-2 / 100000.0)


def call(lib, prefix, handle, dt, gamma, p):
    fn = getattr(lib, f"{prefix}_pressure_control")
    fn.restype = ctypes.c_int
    fn.argtypes = ([ctypes.c_void_p] + [ctypes.c_double] * 7
                   + [ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_double)])
    counts = (ctypes.c_ulonglong * 2)()
    maxima = (ctypes.c_double * 2)()
    rc = fn(handle, dt, gamma, p["density_floor"], p["pressure_floor"], p["light_speed"],
            p["low_density_threshold"], p["low_density_sound_speed"], counts, maxima)
    return rc, tuple(counts), tuple(maxima)


def state_parameters(prob, cfg, gamma):
    """Thresholds at the median density and sound speed of the active cells."""
    ni, nj, nk = cfg.extent
    g, nr, n = cfg.ghost, cfg.radial_cells, cfg.angular_cells
    rho, cs = [], []
    for f in range(gcs_config.FACE_COUNT):
        u = prob.field(f, problem.CONSERVED).reshape(ni, nj, nk, 5)[
            g:g + nr, g:g + n, g:g + n].reshape(-1, 5)
        v2 = (u[:, 1:4] ** 2).sum(1) / u[:, 0] ** 2
        p = (gamma - 1.0) * (u[:, 4] - 0.5 * u[:, 0] * v2)
        rho.append(u[:, 0])
        cs.append(np.sqrt(gamma * p / u[:, 0]))
    rho, cs = np.concatenate(rho), np.concatenate(cs)
    return dict(density_floor=1e-6, pressure_floor=1e-8,
                light_speed=float(np.median(cs)) / 1.5,
                low_density_threshold=float(np.median(rho)),
                low_density_sound_speed=0.5 * float(np.median(cs)))


GAMMA = 5.0 / 3.0


@pytest.mark.parametrize("which", ["driver", "state"])
def test_oracle_pressure_control_bitwise_equal_reference(ref, orc, which):
    cfg = gcs_config.workload("small")
    r = problem.Problem(ref, "refh", cfg)
    r.advance(3)
    o = oracle_from(r, orc, cfg)
    params = DRIVER if which == "driver" else state_parameters(r, cfg, GAMMA)
    dt = r.clock()["dt"]
    rc_r, counts_r, max_r = call(ref, "refh", r._handle, dt, GAMMA, params)
    rc_o, counts_o, max_o = call(orc, "orc", o._handle, dt, GAMMA, params)
    print(f"{which}: reference counts {counts_r} maxima {max_r}", flush=True)
    assert rc_r == rc_o == 0
    assert counts_r == counts_o and max_r == max_o
    if which == "state":
        assert counts_r[0] > 0 and counts_r[1] > 0
    for f in range(gcs_config.FACE_COUNT):
        assert bits_equal(r.field(f, problem.CONSERVED), o.field(f, problem.CONSERVED)), f
    # the reference's argument checks
    bad = dict(params, light_speed=0.0)
    assert call(orc, "orc", o._handle, dt, GAMMA, bad)[0] == -1


@pytest.mark.gpu
def test_gpu_pressure_control_matches_oracle(gpu, orc):
    cfg = gcs_config.workload("small")
    g = gpu.GpuProblem(cfg)
    g.advance(3)
    o = oracle_from(g, orc, cfg)  # the device state, bit for bit
    clk = g.clock()
    orc.orc_set_clock.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_long]
    orc.orc_set_clock(o._handle, clk["time"], clk["dt"], clk["previous_dt"], clk["steps"])
    params = state_parameters(g, cfg, GAMMA)
    dt = g.clock()["dt"]
    rc_g, counts_g, max_g = call(gpu.library(), "gcs_problem", g._handle, dt, GAMMA, params)
    rc_o, counts_o, max_o = call(orc, "orc", o._handle, dt, GAMMA, params)
    print(f"device counts {counts_g} maxima {max_g}; oracle {counts_o} {max_o}", flush=True)
    assert rc_g == rc_o == 0
    assert counts_g == counts_o and counts_g[0] > 0 and counts_g[1] > 0
    assert max_g == max_o
    for f in range(gcs_config.FACE_COUNT):
        assert bits_equal(g.field(f, problem.CONSERVED), o.field(f, problem.CONSERVED)), f
    # the next step takes its dt from the controlled state
    g.advance(1)
    o.advance(1)
    assert abs(g.clock()["dt"] - o.clock()["dt"]) <= 1e-14 * o.clock()["dt"]
    for f in range(gcs_config.FACE_COUNT):
        assert scaled_error(o.field(f, problem.CONSERVED),
                            g.field(f, problem.CONSERVED)) <= 1e-13, f
    assert call(gpu.library(), "gcs_problem", g._handle, dt, 1.0, params)[0] == -1
