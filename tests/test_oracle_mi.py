"""The oracle's ionosphere (MI) coupling against the compiled reference,
bitwise: the dipole mapping helpers, both polar solvers on random inputs,
and a six-face Earth run with MI updates (cubedsphere_mi_dipole.c)."""
import ctypes

import numpy as np
import pytest

from conftest import bits_equal, gcs_config, oracle_from, problem

D = ctypes.c_double
VP = ctypes.c_void_p


class SolverConfig(ctypes.Structure):
    """gamera_mi_solver_config (NO/nonorthogonal_mi.h:54-66)."""
    _fields_ = [("longitude_count", ctypes.c_size_t),
                ("colatitude_count", ctypes.c_size_t),
                ("maximum_colatitude", D), ("ionosphere_radius_m", D),
                ("pedersen_siemens", D), ("hall_siemens", D),
                ("low_latitude_potential_v", D), ("hemisphere", ctypes.c_int),
                ("maximum_iterations", ctypes.c_int),
                ("relative_tolerance", D), ("absolute_tolerance", D)]


class SolverStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("initial_residual", D),
                ("final_residual", D), ("converged", ctypes.c_int)]


def ptr(a):
    return a.ctypes.data_as(VP)


def test_mapping_helpers_match_reference(ref, orc):
    pairs = [("gamera_mi_mapped_colatitude", "orc_mi_mapped_colatitude"),
             ("gamera_mi_dipole_field_ratio", "orc_mi_field_ratio")]
    for r_name, o_name in pairs:
        for lib, name in ((ref, r_name), (orc, o_name)):
            getattr(lib, name).restype = D
            getattr(lib, name).argtypes = [D, D, D]
    ref.gamera_mi_dipole_cos_inclination.restype = D
    ref.gamera_mi_dipole_cos_inclination.argtypes = [D, ctypes.c_int]
    orc.orc_mi_cos_inclination.restype = D
    orc.orc_mi_cos_inclination.argtypes = [D, ctypes.c_int]
    rng = np.random.default_rng(11)
    for _ in range(500):
        theta, radius = rng.uniform(0.0, 1.55), rng.uniform(1.0, 6.0)
        for r_name, o_name in pairs:
            a = getattr(ref, r_name)(theta, radius, 1.0)
            b = getattr(orc, o_name)(theta, radius, 1.0)
            assert a == b or (np.isnan(a) and np.isnan(b))
        for hemi in (-1, 1):
            assert (ref.gamera_mi_dipole_cos_inclination(theta, hemi)
                    == orc.orc_mi_cos_inclination(theta, hemi))
    # closed forms: the ratio is 1 on the ionosphere, cos I = +-1 at the pole
    assert orc.orc_mi_field_ratio(0.3, 1.0, 1.0) == pytest.approx(1.0, rel=1e-15)
    assert orc.orc_mi_cos_inclination(0.0, -1) == -1.0


@pytest.mark.parametrize("tensor", [False, True])
@pytest.mark.parametrize("hemisphere", [-1, 1])
def test_polar_solvers_match_reference(ref, orc, tensor, hemisphere):
    rng = np.random.default_rng(5 + hemisphere + 7 * tensor)
    np_, nt = 32, 14
    cfg = SolverConfig(np_, nt, 0.55, 6.371e6, 10.0, 0.0, 0.0, hemisphere,
                       20000, 1e-10, 0.0)
    fac = rng.normal(size=np_ * nt) * 1e-6
    guess = rng.normal(size=np_ * nt) * 10.0
    pa, pb = guess.copy(), guess.copy()
    sa, sb = SolverStats(), SolverStats()
    if tensor:
        sp = rng.uniform(5.0, 15.0, size=np_ * nt)
        sh = rng.uniform(0.0, 8.0, size=np_ * nt)
        ra = ref.gamera_mi_solve_conductance_tensor(
            ctypes.byref(cfg), ptr(fac), ptr(sp), ptr(sh), ptr(pa), ctypes.byref(sa))
        rb = orc.orc_mi_solve_tensor(
            ctypes.byref(cfg), ptr(fac), ptr(sp), ptr(sh), ptr(pb), ctypes.byref(sb))
        # the residual of the full operator is small (manufactured check)
        out = np.zeros_like(pb)
        assert orc.orc_mi_apply_tensor(ctypes.byref(cfg), ptr(sp), ptr(sh), ptr(pb),
                                       ptr(out)) == 0
    else:
        ra = ref.gamera_mi_solve_constant_pedersen(
            ctypes.byref(cfg), ptr(fac), ptr(pa), ctypes.byref(sa))
        rb = orc.orc_mi_solve_pedersen(
            ctypes.byref(cfg), ptr(fac), ptr(pb), ctypes.byref(sb))
    assert ra == rb == 0
    assert sa.iterations == sb.iterations > 0
    assert sa.converged and sb.converged
    assert bits_equal(pa, pb)


def test_pedersen_operator_inverts_solver(orc):
    """Apply the full constant-Pedersen operator to the solution: the
    interior rows give -R^2 FAC cos(I) / sin^0 (the right-hand side before
    weighting), the pole row the first-ring mean, the outer row zero."""
    np_, nt = 24, 10
    cfg = SolverConfig(np_, nt, 0.5, 6.371e6, 7.0, 0.0, 0.0, -1, 20000,
                       1e-12, 0.0)
    rng = np.random.default_rng(2)
    fac = rng.normal(size=np_ * nt) * 1e-6
    pot = np.zeros(np_ * nt)
    st = SolverStats()
    assert orc.orc_mi_solve_pedersen(ctypes.byref(cfg), ptr(fac), ptr(pot),
                                     ctypes.byref(st)) == 0
    out = np.zeros_like(pot)
    assert orc.orc_mi_apply_pedersen(ctypes.byref(cfg), ptr(pot), ptr(out)) == 0
    out = out.reshape(nt, np_)
    dtheta = 0.5 / (nt - 1)
    theta = np.arange(nt) * dtheta
    cosine = np.cos(theta)
    cos_i = -2.0 * cosine / np.sqrt(1.0 + 3.0 * cosine * cosine)
    expect = -(6.371e6 ** 2) * fac.reshape(nt, np_) * cos_i[:, None]
    scale = np.abs(expect[1:-1]).max()
    assert np.abs(out[1:-1] - expect[1:-1]).max() <= 1e-8 * scale
    assert np.abs(out[0]).max() <= 1e-9 * np.abs(pot).max()
    assert np.abs(out[-1]).max() == 0.0


def test_earth_run_with_mi_bitwise_equal_reference(ref, orc):
    cfg = gcs_config.workload("small_mi")
    r = problem.Problem(ref, "refh", cfg)
    o = oracle_from(r, orc, cfg)
    n = 2 * cfg.mi.azimuth_points * cfg.mi.colatitude_points
    ref.refh_mi_potential.argtypes = [VP, VP, ctypes.c_size_t]
    ref.refh_mi_fac.argtypes = [VP, VP, ctypes.c_size_t]
    ref.refh_mi_updates.argtypes = [VP]
    orc.orc_mi_potential_export.argtypes = [VP, VP, ctypes.c_size_t]
    orc.orc_mi_fac_export.argtypes = [VP, VP, ctypes.c_size_t]
    orc.orc_mi_updates.argtypes = [VP]
    pr, po, fr, fo = (np.zeros(n) for _ in range(4))
    for step in range(45):  # two MI updates (t = 0 and t = 10 s)
        r.advance(1)
        o.advance(1)
        for fid in problem.STATE_FIELDS:
            for f in range(gcs_config.FACE_COUNT):
                assert bits_equal(r.field(f, fid), o.field(f, fid)), (step, fid, f)
    assert ref.refh_mi_updates(r._handle) == orc.orc_mi_updates(o._handle) == 2
    assert ref.refh_mi_potential(r._handle, ptr(pr), n) == 0
    assert orc.orc_mi_potential_export(o._handle, ptr(po), n) == 0
    assert ref.refh_mi_fac(r._handle, ptr(fr), n) == 0
    assert orc.orc_mi_fac_export(o._handle, ptr(fo), n) == 0
    assert bits_equal(pr, po) and bits_equal(fr, fo)
    assert np.abs(pr).max() > 1.0  # the coupling is active, not trivially zero
