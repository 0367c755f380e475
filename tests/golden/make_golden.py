"""Write tests/golden/pointwise.npz: input/output vectors of the reference's
pointwise step functions, computed by the compiled reference library
(oracle/_ref/libgamera_ref.so, built by `make -C oracle ref`).

The reference ships no data fixtures (its tests are self-checking), so these
vectors are generated here from the reference's own code on seeded random
inputs and committed; tests/test_oracle_units.py checks the oracle's
restatements against them bitwise.

    python tests/golden/make_golden.py
"""
import ctypes
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
import refabi as R  # noqa: E402

N = 64


def vec(v):
    return R.Vec3((ctypes.c_double * 3)(*v))


def prim(w):
    return R.Primitive(w[0], vec(w[1:4]), w[4])


def arr(x):
    return (ctypes.c_double * len(x))(*x)


def frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t = rng.normal(size=3)
    t -= t.dot(n) * n
    t /= np.linalg.norm(t)
    return n, np.cross(t, n), t


def face_geometry(rng):
    n, t1, t2 = frame(rng)
    area = rng.uniform(0.5, 2.0)
    g = R.FaceGeometry()
    g.area = area
    g.area_vector = vec(area * n)
    g.centroid = vec(rng.normal(size=3))
    g.normal, g.tangent1, g.tangent2 = vec(n), vec(t1), vec(t2)
    return g


def edge_geometry(rng):
    n, t1, t2 = frame(rng)
    g = R.EdgeGeometry()
    g.length = rng.uniform(0.1, 1.0)
    g.normal, g.tangent1, g.tangent2 = vec(n), vec(t1), vec(t2)
    return g


def state(rng):
    return [rng.uniform(0.5, 5.0), *rng.normal(0, 3, size=3),
            rng.uniform(0.05, 2.0)]


def cons(rng):
    w = state(rng)
    v2 = sum(x * x for x in w[1:4])
    return [w[0], *(w[0] * np.array(w[1:4])), w[4] / (2.0 / 3.0) + 0.5 * w[0] * v2]


def make_series(rng, linear):
    times = (ctypes.c_double * 3)(0.0, 0.7, 1.9)
    states = (R.WindState * 3)()
    for s in range(3):
        states[s] = R.WindState(5.0 + s, 0.4 + 0.1 * s,
                                (ctypes.c_double * 3)(4.0 + 0.5 * s, 0.1 * s, 0.0),
                                (ctypes.c_double * 3)(0.2, -0.1, -1.09 + s))
    ser = R.WindSeries(3, times, states, (ctypes.c_double * 3)(-30.0, 0.0, 0.0),
                       0.25, 0.1, -0.2, 0.0, 0, linear)
    return ser, times, states


def generate(ref):
    rng = np.random.default_rng(20261202)
    out = {}
    f = R.bind(ref, "ref")
    # Up7/PDM
    vols, qs, res = [], [], []
    for n in range(N):
        vol = rng.uniform(0.5, 2.0, size=8)
        q = rng.normal(size=8) if n % 3 else np.full(8, rng.normal())
        l, r = ctypes.c_double(), ctypes.c_double()
        st = f["up7_pdm"](arr(vol), arr(q), 1.0, ctypes.byref(l), ctypes.byref(r))
        vols.append(vol), qs.append(q), res.append([st, l.value, r.value])
    out["up7_vol"], out["up7_q"], out["up7_out"] = map(np.array, (vols, qs, res))
    # conserved -> primitive and the AB predictor
    us, uos, ratios, prims, preds = [], [], [], [], []
    for n in range(N):
        u, uo, ratio = cons(rng), cons(rng), rng.uniform(0.0, 0.7)
        w = R.Primitive()
        st = f["cons_to_prim"](arr(u), 5.0 / 3.0, 1e-6, 1e-8, ctypes.byref(w))
        prims.append([st, w.density, *w.velocity.value, w.pressure])
        p = (ctypes.c_double * 5)()
        st = f["predict_cell"](arr(uo), arr(u), ratio, 5.0 / 3.0, 1e-6, 1e-8, p)
        preds.append([st, *p])
        us.append(u), uos.append(uo), ratios.append(ratio)
    out["c2p_u"], out["c2p_out"] = np.array(us), np.array(prims)
    out["pred_old"], out["pred_ratio"], out["pred_out"] = (
        np.array(uos), np.array(ratios), np.array(preds))
    # kinetic fluid + Maxwell fluxes and HOGS on the same interfaces
    fin, fout, mout, hout = [], [], [], []
    for n in range(N):
        wl, wr = state(rng), state(rng) if n % 4 else None
        wr = wr if wr is not None else list(wl)
        g = face_geometry(rng)
        bl, br = rng.normal(0, 2, 3), rng.normal(0, 2, 3)
        bn, b0, b0n = rng.normal(), rng.normal(0, 3, 3), rng.normal()
        pr = (R.Primitive * 2)(prim(wl), prim(wr))
        fl = R.FluidFlux()
        st = f["fluid_flux"](pr, 5.0 / 3.0, ctypes.byref(g), ctypes.byref(fl))
        mx = R.MaxwellFlux()
        un = arr(list(fl.normal_velocity))
        st2 = f["maxwell_flux"](pr, (R.Vec3 * 2)(vec(bl), vec(br)), bn,
                                ctypes.byref(g), True, vec(b0), b0n, True, 10.0,
                                un, ctypes.byref(mx))
        fout.append([st, *fl.conserved, *fl.conserved_jump,
                     *fl.velocity_jump.value, *fl.normal_velocity])
        mout.append([st2, *mx.momentum.value, mx.alfven_diffusion_speed,
                     mx.magnetic_pressure_sum])
        st3 = f["hogs"](ctypes.byref(fl), ctypes.byref(mx), 0.25, 0.25, True, 10.0)
        hout.append([st3, *fl.conserved, *mx.momentum.value])
        fin.append([*wl, *wr, g.area, *g.area_vector.value, *g.centroid.value,
                    *g.normal.value, *g.tangent1.value, *g.tangent2.value,
                    *bl, *br, bn, *b0, b0n])
    out["flux_in"], out["fluid_out"], out["maxwell_out"], out["hogs_out"] = map(
        np.array, (fin, fout, mout, hout))
    # Boris correction
    bin_, bout = [], []
    for n in range(N):
        u, uo = cons(rng), cons(rng)
        bnew, bold = rng.normal(0, 20 if n % 2 else 1e-7, 3), rng.normal(0, 20, 3)
        rate, mr = rng.normal(size=5), rng.normal(size=3)
        dt = rng.uniform(1e-3, 1e-2)
        x = arr(u)
        st = f["boris"](x, arr(uo), vec(bnew), vec(bold), arr(rate), vec(mr), 10.0,
                        dt, 5.0 / 3.0, 1e-6, 1e-8)
        bin_.append([*u, *uo, *bnew, *bold, *rate, *mr, dt])
        bout.append([st, *x])
    out["boris_in"], out["boris_out"] = np.array(bin_), np.array(bout)
    # edge EMF, 2x2 edge solve, face normal at edge
    ein, eout = [], []
    for n in range(N):
        g = edge_geometry(rng)
        a = rng.normal(size=6)
        rho = rng.uniform(0.5, 5.0)
        dt = rng.uniform(1e-3, 1e-1)
        emf, speed = ctypes.c_double(), ctypes.c_double()
        st = f["edge_emf"](a[0], a[1], a[2], a[3], a[4], rho, ctypes.byref(g),
                           0.5, True, 10.0, 0.3, dt, ctypes.byref(emf),
                           ctypes.byref(speed))
        n1, n2 = rng.normal(size=2), rng.normal(size=2)
        sol = (ctypes.c_double * 2)()
        st2 = f["edge_solve"](arr(n1), arr(n2), a[5], a[0], sol)
        areas = rng.uniform(0.5, 2.0, 8)
        normals = [vec(frame(rng)[0] * 0.1 + np.array([0.0, 0.0, 1.0]))
                   for _ in range(8)]
        tn = (ctypes.c_double * 2)()
        st3 = f["normal_at_edge"](arr(areas), (R.Vec3 * 8)(*normals),
                                  ctypes.byref(g), tn)
        ein.append([g.length, *g.normal.value, *g.tangent1.value,
                    *g.tangent2.value, *a, rho, dt, *n1, *n2, *areas,
                    *[c for v in normals for c in v.value]])
        eout.append([st, emf.value, speed.value, st2, *sol, st3, *tn])
    out["edge_in"], out["edge_out"] = np.array(ein), np.array(eout)
    # calibrated recovery on real cells of a small cube problem
    cfg_mod = importlib.import_module("gamera-cube_amd.config")
    prob_mod = importlib.import_module("gamera-cube_amd.problem")
    p = prob_mod.Problem(ref, "refh", cfg_mod.workload("tiny"))
    cells = p.field(0, prob_mod.CELL_GEOMETRY).reshape(-1, 103)
    pick = rng.choice(cells.shape[0], N, replace=False)
    rin, rout = [], []
    for n, c in enumerate(pick):
        geo = R.CellGeometry.from_buffer_copy(cells[c].tobytes())
        flux = rng.normal(size=6)
        b = R.Vec3()
        net = ctypes.c_double()
        st = f["recover"](ctypes.byref(geo), arr(flux), ctypes.byref(b),
                          ctypes.byref(net))
        rin.append([*cells[c], *flux])
        rout.append([st, *b.value])
    p.close()
    out["rec_in"], out["rec_out"] = np.array(rin), np.array(rout)
    # solar wind sampling and weight (linear and piecewise-constant series)
    win, wout = [], []
    for n in range(N):
        linear = n % 2
        ser, _t, _s = make_series(rng, linear)
        x = rng.normal(0, 20, 3)
        t = rng.uniform(-1.0, 4.0)
        s = R.WindState()
        delay = ctypes.c_double()
        st = f["wind_sample"](ctypes.byref(ser), arr(x), t, ctypes.byref(s),
                              ctypes.byref(delay))
        w = ctypes.c_double()
        st2 = f["wind_weight"](ctypes.byref(ser), arr(x), t, ctypes.byref(w))
        win.append([linear, *x, t])
        wout.append([st, s.density, s.pressure, *s.velocity, *s.magnetic, st2,
                     w.value])
    out["wind_in"], out["wind_out"] = np.array(win), np.array(wout)
    # cube inverse and edge neighbours
    iin, iout = [], []
    for n in range(N):
        d = rng.normal(size=3)
        face, a, b = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        st = f["inverse"](vec(d), ctypes.byref(face), ctypes.byref(a),
                          ctypes.byref(b))
        iin.append(d)
        iout.append([st, face.value, a.value, b.value])
    out["inv_in"], out["inv_out"] = np.array(iin), np.array(iout)

    class Neighbor(ctypes.Structure):
        _fields_ = [("face", ctypes.c_int), ("edge", ctypes.c_int),
                    ("reversed", ctypes.c_int)]
    fn = ref.gamera_cs_edge_neighbor
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Neighbor)]
    nb = []
    for face in range(6):
        for edge in range(4):
            x = Neighbor()
            st = fn(face, edge, ctypes.byref(x))
            nb.append([face, edge, st, x.face, x.edge, x.reversed])
    out["neighbor"] = np.array(nb)
    # emergency interface diffusion (nonorthogonal_flux.c:206-241): interface
    # speeds drawn around the 1.5 c threshold, light speeds 10 and 2, Boris
    # off, an exactly-at-threshold pair, a non-physical interface and a
    # non-finite cell value part-way through the variables.
    ein, eout = [], []
    for n in range(N):
        c = 10.0 if n % 2 else 2.0
        boris = n % 11 != 5
        wl, wr = state(rng), state(rng)
        scale = rng.uniform(0.5, 2.0) * 1.5 * c
        for w in (wl, wr):
            v = np.array(w[1:4])
            w[1:4] = list(v / np.linalg.norm(v) * scale * rng.uniform(0.8, 1.2))
        if n % 13 == 3:      # sqrt(0.5 (225 + 225)) == 15 == 1.5 c exactly
            c = 10.0
            wl[1:4], wr[1:4] = [15.0, 0.0, 0.0], [0.0, -15.0, 0.0]
        if n % 17 == 7:
            wr[0] = -1.0     # rejected interface state
        lo, up = rng.normal(0, 3, 5), rng.normal(0, 3, 5)
        if n % 19 == 9:
            up[3] = np.inf   # the first three variables are already updated
        fl = R.FluidFlux()
        flux0 = rng.normal(size=5)
        fl.conserved[:] = list(flux0)
        applied = ctypes.c_bool(False)
        st = f["emergency"](ctypes.byref(fl), (R.Primitive * 2)(prim(wl), prim(wr)),
                            arr(lo), arr(up), boris, c, ctypes.byref(applied))
        ein.append([c, float(boris), *wl, *wr, *lo, *up, *flux0])
        eout.append([st, float(applied.value), *fl.conserved])
    out["emerg_in"], out["emerg_out"] = np.array(ein), np.array(eout)
    return out


def main():
    lib = os.path.join(ROOT, "oracle", "_ref", "libgamera_ref.so")
    ref = ctypes.CDLL(lib)
    data = generate(ref)
    path = os.path.join(HERE, "pointwise.npz")
    np.savez_compressed(path, **data)
    print(path, sum(v.size for v in data.values()), "values")


if __name__ == "__main__":
    main()
