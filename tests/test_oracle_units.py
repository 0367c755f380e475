"""The oracle's pointwise functions against the reference's own outputs
(golden vectors made by tests/golden/make_golden.py from the compiled
reference) and against closed forms that pin them independently."""
import ctypes
import math
import os

import numpy as np
import pytest

import refabi as R
from conftest import ROOT, bits_equal

GOLDEN = np.load(os.path.join(ROOT, "tests", "golden", "pointwise.npz"))


def arr(x):
    return (ctypes.c_double * len(x))(*x)


def vec(v):
    return R.Vec3((ctypes.c_double * 3)(*v))


@pytest.fixture(scope="module")
def fo(orc):
    return R.bind(orc, "orc")


def test_up7_pdm_matches_reference(fo):
    for vol, q, want in zip(GOLDEN["up7_vol"], GOLDEN["up7_q"], GOLDEN["up7_out"]):
        l, r = ctypes.c_double(), ctypes.c_double()
        st = fo["up7_pdm"](arr(vol), arr(q), 1.0, ctypes.byref(l), ctypes.byref(r))
        assert bits_equal([st, l.value, r.value], want)


def test_primitive_and_predictor_match_reference(fo):
    for u, want in zip(GOLDEN["c2p_u"], GOLDEN["c2p_out"]):
        w = R.Primitive()
        st = fo["cons_to_prim"](arr(u), 5.0 / 3.0, 1e-6, 1e-8, ctypes.byref(w))
        assert bits_equal([st, w.density, *w.velocity.value, w.pressure], want)
    for uo, u, r, want in zip(GOLDEN["pred_old"], GOLDEN["c2p_u"],
                              GOLDEN["pred_ratio"], GOLDEN["pred_out"]):
        p = (ctypes.c_double * 5)()
        st = fo["predict_cell"](arr(uo), arr(u), r, 5.0 / 3.0, 1e-6, 1e-8, p)
        assert bits_equal([st, *p], want)


def unpack_flux_case(row):
    wl, wr = row[0:5], row[5:10]
    g = R.FaceGeometry()
    g.area = row[10]
    g.area_vector, g.centroid = vec(row[11:14]), vec(row[14:17])
    g.normal, g.tangent1, g.tangent2 = vec(row[17:20]), vec(row[20:23]), vec(row[23:26])
    bl, br, bn, b0, b0n = row[26:29], row[29:32], row[32], row[33:36], row[36]
    pr = (R.Primitive * 2)(R.Primitive(wl[0], vec(wl[1:4]), wl[4]),
                           R.Primitive(wr[0], vec(wr[1:4]), wr[4]))
    return pr, g, bl, br, bn, b0, b0n


def test_kinetic_fluxes_and_hogs_match_reference(fo):
    for row, fw, mw, hw in zip(GOLDEN["flux_in"], GOLDEN["fluid_out"],
                               GOLDEN["maxwell_out"], GOLDEN["hogs_out"]):
        pr, g, bl, br, bn, b0, b0n = unpack_flux_case(row)
        fl = R.FluidFlux()
        st = fo["fluid_flux"](pr, 5.0 / 3.0, ctypes.byref(g), ctypes.byref(fl))
        assert bits_equal([st, *fl.conserved, *fl.conserved_jump,
                           *fl.velocity_jump.value, *fl.normal_velocity], fw)
        mx = R.MaxwellFlux()
        st = fo["maxwell_flux"](pr, (R.Vec3 * 2)(vec(bl), vec(br)), bn,
                                ctypes.byref(g), True, vec(b0), b0n, True, 10.0,
                                arr(list(fl.normal_velocity)), ctypes.byref(mx))
        assert bits_equal([st, *mx.momentum.value, mx.alfven_diffusion_speed,
                           mx.magnetic_pressure_sum], mw)
        st = fo["hogs"](ctypes.byref(fl), ctypes.byref(mx), 0.25, 0.25, True, 10.0)
        assert bits_equal([st, *fl.conserved, *mx.momentum.value], hw)


def test_emergency_diffusion_matches_reference(fo):
    """orc_emergency_diffusion against nonorthogonal_flux.c:206-241 on
    interfaces around the 1.5 c threshold (one exactly at it), Boris off,
    a rejected interface and a non-finite cell value."""
    applied_cases = 0
    for row, want in zip(GOLDEN["emerg_in"], GOLDEN["emerg_out"]):
        c, boris = row[0], bool(row[1])
        pr = (R.Primitive * 2)(R.Primitive(row[2], vec(row[3:6]), row[6]),
                               R.Primitive(row[7], vec(row[8:11]), row[11]))
        fl = R.FluidFlux()
        fl.conserved[:] = list(row[22:27])
        applied = ctypes.c_bool(True)
        st = fo["emergency"](ctypes.byref(fl), pr, arr(row[12:17]), arr(row[17:22]),
                             boris, c, ctypes.byref(applied))
        assert bits_equal([st, float(applied.value), *fl.conserved], want)
        applied_cases += int(applied.value)
        if st == 0 and applied.value:
            # the applied branch subtracts c (upper - lower), nothing else
            expect = row[22:27] - c * (row[17:22] - row[12:17])
            assert bits_equal(list(fl.conserved), expect)
    assert applied_cases >= 16


def test_boris_matches_reference(fo):
    for row, want in zip(GOLDEN["boris_in"], GOLDEN["boris_out"]):
        u = arr(row[0:5])
        st = fo["boris"](u, arr(row[5:10]), vec(row[10:13]), vec(row[13:16]),
                         arr(row[16:21]), vec(row[21:24]), 10.0, row[24],
                         5.0 / 3.0, 1e-6, 1e-8)
        assert bits_equal([st, *u], want)


def test_edge_operators_match_reference(fo):
    for row, want in zip(GOLDEN["edge_in"], GOLDEN["edge_out"]):
        g = R.EdgeGeometry()
        g.length = row[0]
        g.normal, g.tangent1, g.tangent2 = vec(row[1:4]), vec(row[4:7]), vec(row[7:10])
        a, rho, dt = row[10:16], row[16], row[17]
        emf, speed = ctypes.c_double(), ctypes.c_double()
        st = fo["edge_emf"](a[0], a[1], a[2], a[3], a[4], rho, ctypes.byref(g), 0.5,
                            True, 10.0, 0.3, dt, ctypes.byref(emf), ctypes.byref(speed))
        sol = (ctypes.c_double * 2)()
        st2 = fo["edge_solve"](arr(row[18:20]), arr(row[20:22]), a[5], a[0], sol)
        normals = (R.Vec3 * 8)(*[vec(row[30 + 3 * n:33 + 3 * n]) for n in range(8)])
        tn = (ctypes.c_double * 2)()
        st3 = fo["normal_at_edge"](arr(row[22:30]), normals, ctypes.byref(g), tn)
        got = [st, emf.value, speed.value, st2, *sol, st3, *tn]
        # outputs of failed calls are not defined
        if st2 != 0:
            got[4:6] = want[4:6]
        assert bits_equal(got, want)


def test_calibrated_recovery_matches_reference(fo):
    for row, want in zip(GOLDEN["rec_in"], GOLDEN["rec_out"]):
        geo = R.CellGeometry.from_buffer_copy(np.ascontiguousarray(row[:103]).tobytes())
        b = R.Vec3()
        st = fo["recover"](ctypes.byref(geo), arr(row[103:109]), ctypes.byref(b))
        assert bits_equal([st, *b.value], want)


def test_solar_wind_matches_reference(fo):
    for row, want in zip(GOLDEN["wind_in"], GOLDEN["wind_out"]):
        times = (ctypes.c_double * 3)(0.0, 0.7, 1.9)
        states = (R.WindState * 3)()
        for s in range(3):
            states[s] = R.WindState(5.0 + s, 0.4 + 0.1 * s,
                                    (ctypes.c_double * 3)(4.0 + 0.5 * s, 0.1 * s, 0.0),
                                    (ctypes.c_double * 3)(0.2, -0.1, -1.09 + s))
        ser = R.WindSeries(3, times, states, (ctypes.c_double * 3)(-30.0, 0.0, 0.0),
                           0.25, 0.1, -0.2, 0.0, 0, int(row[0]))
        s = R.WindState()
        st = fo["wind_sample"](ctypes.byref(ser), arr(row[1:4]), row[4],
                               ctypes.byref(s))
        w = ctypes.c_double()
        st2 = fo["wind_weight"](ctypes.byref(ser), arr(row[1:4]), row[4],
                                ctypes.byref(w))
        assert bits_equal([st, s.density, s.pressure, *s.velocity, *s.magnetic,
                           st2, w.value], want)


def test_cube_topology_matches_reference(fo):
    for d, want in zip(GOLDEN["inv_in"], GOLDEN["inv_out"]):
        face, a, b = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        st = fo["inverse"](vec(d), ctypes.byref(face), ctypes.byref(a), ctypes.byref(b))
        assert bits_equal([st, face.value, a.value, b.value], want)
    for face, edge, st_ref, nf_ref, ne_ref, rev_ref in GOLDEN["neighbor"].astype(int):
        nf, ne, rev = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        st = fo["neighbor"](face, edge, ctypes.byref(nf), ctypes.byref(ne),
                            ctypes.byref(rev))
        assert (st, nf.value, ne.value, rev.value) == (st_ref, nf_ref, ne_ref, rev_ref)


# ---- closed forms, independent of the reference's outputs -------------------

def test_up7_pdm_closed_forms(fo):
    rng = np.random.default_rng(1)
    for _ in range(50):
        vol = rng.uniform(0.5, 2.0, 8)
        c = rng.normal()
        l, r = ctypes.c_double(), ctypes.c_double()
        fo["up7_pdm"](arr(vol), arr([c] * 8), 1.0, ctypes.byref(l), ctypes.byref(r))
        assert l.value == c and r.value == c          # constants are exact
        q = rng.normal() + rng.normal() * np.arange(8)  # linear, equal volumes
        fo["up7_pdm"](arr([1.0] * 8), arr(q), 1.0, ctypes.byref(l), ctypes.byref(r))
        mid = 0.5 * (q[3] + q[4])
        assert abs(l.value - mid) <= 1e-13 * (1 + abs(mid))
        assert abs(r.value - mid) <= 1e-13 * (1 + abs(mid))
        q = rng.normal(size=8)                        # PDM keeps within neighbours
        fo["up7_pdm"](arr(vol), arr(q), 1.0, ctypes.byref(l), ctypes.byref(r))
        lo, hi = min(q[3], q[4]), max(q[3], q[4])
        assert lo - 1e-15 <= l.value <= hi + 1e-15
        assert lo - 1e-15 <= r.value <= hi + 1e-15


def test_equal_states_give_exact_euler_and_maxwell_flux(fo):
    rng = np.random.default_rng(2)
    gamma = 5.0 / 3.0
    for _ in range(50):
        rho, p = rng.uniform(0.5, 5), rng.uniform(0.1, 3)
        v = rng.normal(0, 2, 3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t2 = np.cross(n, rng.normal(size=3))
        t2 /= np.linalg.norm(t2)
        t1 = np.cross(t2, n)
        g = R.FaceGeometry()
        g.area, g.normal, g.tangent1, g.tangent2 = 1.0, vec(n), vec(t1), vec(t2)
        pr = (R.Primitive * 2)(R.Primitive(rho, vec(v), p), R.Primitive(rho, vec(v), p))
        fl = R.FluidFlux()
        assert fo["fluid_flux"](pr, gamma, ctypes.byref(g), ctypes.byref(fl)) == 0
        un = v.dot(n)
        e = 0.5 * rho * v.dot(v) + p / (gamma - 1)
        exact = [rho * un, *(rho * v * un + p * n), (e + p) * un]
        scale = 1 + np.max(np.abs(exact))
        assert np.max(np.abs(np.array(fl.conserved[:]) - exact)) <= 1e-13 * scale
        b = rng.normal(0, 2, 3)
        bn = b.dot(n)
        mx = R.MaxwellFlux()
        zero = vec([0.0, 0.0, 0.0])
        assert fo["maxwell_flux"](pr, (R.Vec3 * 2)(vec(b), vec(b)), bn, ctypes.byref(g),
                                  True, zero, 0.0, False, 10.0,
                                  arr(list(fl.normal_velocity)), ctypes.byref(mx)) == 0
        exact = 0.5 * b.dot(b) * n - b * bn
        assert np.max(np.abs(np.array(mx.momentum.value[:]) - exact)) <= 1e-13 * (1 + b.dot(b))


def test_recovery_reproduces_uniform_field(fo):
    for row in GOLDEN["rec_in"][:32]:
        geo = R.CellGeometry.from_buffer_copy(np.ascontiguousarray(row[:103]).tobytes())
        bfield = np.array([0.3, -1.7, 2.2])
        flux = []
        for d in range(3):
            for s in range(2):
                av = np.array(geo.face[2 * d + s].area_vector.value[:])
                flux.append(float(bfield.dot(av)))
        b = R.Vec3()
        assert fo["recover"](ctypes.byref(geo), arr(flux), ctypes.byref(b)) == 0
        assert np.max(np.abs(np.array(b.value[:]) - bfield)) <= 1e-12


def test_wind_weight_closed_form(fo):
    times = (ctypes.c_double * 1)(0.0)
    states = (R.WindState * 1)(R.WindState(5.0, 0.4, (ctypes.c_double * 3)(4.0, 0, 0),
                                           (ctypes.c_double * 3)(0, 0, -1.0)))
    ser = R.WindSeries(1, times, states, (ctypes.c_double * 3)(0, 0, 0), 0, 0, 0, 0, 0, 1)
    w = ctypes.c_double()
    for x, want in (([-5.0, 0, 0], 1.0), ([0.0, 5.0, 0], 1.0), ([5.0, 0, 0], 0.0)):
        assert fo["wind_weight"](ctypes.byref(ser), arr(x), 0.0, ctypes.byref(w)) == 0
        assert w.value == want
    # 135 degrees from upstream: exp(1 - 1/(1 - 0.25)) = exp(-1/3)
    fo["wind_weight"](ctypes.byref(ser), arr([1.0, 1.0, 0.0]), 0.0, ctypes.byref(w))
    assert abs(w.value - math.exp(-1.0 / 3.0)) <= 1e-15
