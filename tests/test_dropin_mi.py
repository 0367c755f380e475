"""The dipole ionosphere through the reference's interface
(include/gamera_cs_mi_dropin.h, cubedsphere/cubedsphere_mi_dipole.h:30-98),
against the reference's own gamera_cs_mi_dipole_* on the same problems.

Two copies of one MI-coupled reference problem run side by side: one with
the reference's step and ionosphere, the other with libgcs.so's step and an
ionosphere state made by libgcs.so's gamera_cs_mi_dipole_create on the same
grid and storage (oracle/ref_harness.c refh_use_mi), so its Earth context's
`mi` is the library's own and every step stays on the device. Both solves
run to 1e-13 here (the device's Jacobi-PCG reduces in another order)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REF_LIB, gcs_config, problem, scaled_error, state_relative_errors

pytestmark = pytest.mark.gpu

VP = ctypes.c_void_p
STEP_TOL = 1e-13


class Vec3(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double * 3)]


class SolverStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("initial_residual", ctypes.c_double),
                ("final_residual", ctypes.c_double), ("converged", ctypes.c_int)]


@pytest.fixture(scope="module")
def libs(gpu):
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    ref = ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_GLOBAL)
    for name in ("refh_grid", "refh_storage", "refh_background", "refh_startup"):
        getattr(ref, name).restype = VP
        getattr(ref, name).argtypes = [VP]
    ref.refh_set_step.argtypes = [VP, VP]
    ref.refh_use_mi.argtypes = [VP, VP, VP]
    gcs = gpu.library()
    for lib in (ref, gcs):
        lib.gamera_cs_mi_dipole_potential.argtypes = [VP, Vec3, ctypes.POINTER(ctypes.c_double)]
        lib.gamera_cs_mi_dipole_ghost_velocity.argtypes = [VP, Vec3, Vec3, ctypes.POINTER(Vec3)]
        lib.gamera_cs_mi_dipole_export_potential.argtypes = [
            VP, VP, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_double)]
        lib.gamera_cs_mi_dipole_import_potential.argtypes = [
            VP, VP, ctypes.c_size_t, ctypes.c_double, ctypes.c_double]
        lib.gamera_cs_mi_dipole_create_potential_view.argtypes = [VP, VP, ctypes.POINTER(VP)]
        lib.gamera_cs_mi_dipole_export_shell_current.argtypes = [VP, VP, VP, VP, ctypes.c_size_t]
        lib.gamera_cs_mi_dipole_stats.argtypes = [VP] + [ctypes.POINTER(ctypes.c_double)] * 4 + [
            ctypes.POINTER(SolverStats)]
        lib.gamera_cs_mi_dipole_apply_inner_emf.argtypes = [VP, VP, VP, ctypes.c_size_t]
        lib.gamera_cs_mi_dipole_destroy.argtypes = [VP]
        lib.gamera_cs_mi_dipole_ready.argtypes = [VP]
    gcs.gamera_cs_mi_dipole_create.argtypes = [VP, VP, VP, VP, VP, ctypes.c_size_t, VP,
                                               ctypes.POINTER(VP)]
    gcs.gcs_dropin_host_round_trips.restype = ctypes.c_long
    gcs.gcs_dropin_release.argtypes = [VP, VP]
    ref.refh_mi.restype = VP
    ref.refh_mi.argtypes = [VP]
    return ref, gcs


def mi_config(cfg):
    cfg.mi.relative_tolerance = 1e-13
    return cfg


def export_potential(lib, mi, cfg):
    n = 2 * cfg.mi.azimuth_points * cfg.mi.colatitude_points
    out = np.zeros(n)
    epoch, nxt = ctypes.c_double(), ctypes.c_double()
    assert lib.gamera_cs_mi_dipole_export_potential(mi, out.ctypes.data, n, ctypes.byref(epoch),
                                                    ctypes.byref(nxt)) == 0
    return out, epoch.value, nxt.value


def create_gcs_mi(ref, gcs, p, cfg):
    active = (ctypes.c_size_t * 3)(cfg.radial_cells, cfg.angular_cells, cfg.angular_cells)
    h = VP()
    rc = gcs.gamera_cs_mi_dipole_create(ref.refh_grid(p._handle), ref.refh_storage(p._handle),
                                        ref.refh_background(p._handle),
                                        ref.refh_startup(p._handle), active, cfg.ghost,
                                        ctypes.byref(cfg.mi), ctypes.byref(h))
    assert rc == 0 and h.value
    return h


def test_mi_coupled_steps_through_the_reference_interface(libs, tmp_path):
    """45 driver steps (MI updates at t = 0 and 10 s) of a small Earth
    problem: the drop-in step with libgcs.so's ionosphere against the
    reference's step and ionosphere; no host round trip."""
    ref, gcs = libs
    cfg = mi_config(gcs_config.workload("small_mi"))
    a = problem.Problem(ref, "refh", cfg)
    b = problem.Problem(ref, "refh", cfg)
    ref.refh_set_step(b._handle, ctypes.cast(gcs.gamera_cs_advance_six_with_background_planned,
                                             VP).value)
    h = create_gcs_mi(ref, gcs, b, cfg)
    ref.refh_use_mi(b._handle, h, ctypes.cast(gcs.gamera_cs_mi_dipole_maybe_update, VP).value)
    rounds = gcs.gcs_dropin_host_round_trips()
    try:
        worst = 0.0
        for s in range(45):
            a.advance(1)
            b.advance(1)
            for fid in problem.STATE_FIELDS:
                for f in range(gcs_config.FACE_COUNT):
                    worst = max(worst, scaled_error(a.field(f, fid), b.field(f, fid)))
            assert worst <= STEP_TOL * (s + 1), (s, worst)
        rel = max(state_relative_errors(a, b).values())
        ref_mi = ref.refh_mi(a._handle)
        pa, ea, na = export_potential(ref, ref_mi, cfg)
        pb, eb, nb = export_potential(gcs, h, cfg)
        pot_err = np.abs(pa - pb).max() / np.abs(pa).max()
        sa = [ctypes.c_double() for _ in range(4)]
        sb = [ctypes.c_double() for _ in range(4)]
        ta, tb = (SolverStats * 2)(), (SolverStats * 2)()
        assert ref.gamera_cs_mi_dipole_stats(ref_mi, *[ctypes.byref(x) for x in sa], ta) == 0
        assert gcs.gamera_cs_mi_dipole_stats(h, *[ctypes.byref(x) for x in sb], tb) == 0
        print(f"MI drop-in: 45 steps scaled {worst:.2e} relative {rel:.2e}; potential rel "
              f"{pot_err:.2e}; FAC range ref [{sa[0].value:.6e}, {sa[1].value:.6e}] gpu "
              f"[{sb[0].value:.6e}, {sb[1].value:.6e}]; iterations ref "
              f"({ta[0].iterations}, {ta[1].iterations}) gpu ({tb[0].iterations}, "
              f"{tb[1].iterations}); host round trips "
              f"{gcs.gcs_dropin_host_round_trips() - rounds}", flush=True)
        assert rel <= STEP_TOL * 45
        assert (ea, na) == (eb, nb)
        assert pot_err <= 1e-10
        for x, y in zip(sa, sb):
            assert abs(x.value - y.value) <= 1e-9 * max(abs(sa[0].value), abs(sa[1].value),
                                                        abs(sa[2].value), abs(sa[3].value))
        assert all(tb[k].converged for k in range(2))
        assert gcs.gcs_dropin_host_round_trips() == rounds
        # the shell current of the last update, [face][j][k]
        n = 6 * cfg.angular_cells ** 2
        out = {lib: [np.zeros(n) for _ in range(3)] for lib in ("ref", "gcs")}
        for key, lib, mi in (("ref", ref, ref_mi), ("gcs", gcs, h)):
            o = out[key]
            assert lib.gamera_cs_mi_dipole_export_shell_current(
                mi, o[0].ctypes.data, o[1].ctypes.data, o[2].ctypes.data, n) == 0
        for q in range(3):
            x, y = out["ref"][q], out["gcs"][q]
            assert np.abs(x - y).max() <= 1e-9 * np.abs(x).max(), q
        # the CSV files (cubedsphere_mi_dipole.c:620-690): same headers,
        # labels, grid coordinates and clock; values to the solve tolerance
        files = {}
        for key, lib, mi in (("ref", ref, ref_mi), ("gcs", gcs, h)):
            prefix = str(tmp_path / key).encode()
            lib.gamera_cs_mi_dipole_write_csv.argtypes = [VP, ctypes.c_char_p]
            lib.gamera_cs_mi_dipole_write_shell_current_csv.argtypes = [VP, ctypes.c_char_p]
            assert lib.gamera_cs_mi_dipole_write_csv(mi, prefix) == 0
            assert lib.gamera_cs_mi_dipole_write_shell_current_csv(mi, prefix) == 0
            files[key] = {ext: (tmp_path / f"{key}.{ext}").read_text().splitlines()
                          for ext in ("mi_north.csv", "mi_south.csv", "shell_j.csv")}
        for ext, exact_columns in (("mi_north.csv", 4), ("mi_south.csv", 4), ("shell_j.csv", 3)):
            fa, fb = files["ref"][ext], files["gcs"][ext]
            assert fa[0] == fb[0] and len(fa) == len(fb), ext
            ra = [line.split(",") for line in fa[1:]]
            rb = [line.split(",") for line in fb[1:]]
            assert all(x[:exact_columns] == y[:exact_columns] for x, y in zip(ra, rb)), ext
            va = np.array([[float(v) for v in x[exact_columns:]] for x in ra])
            vb = np.array([[float(v) for v in x[exact_columns:]] for x in rb])
            scale = np.maximum(np.abs(va).max(axis=0), 1e-300)
            assert (np.abs(va - vb).max(axis=0) <= 1e-9 * scale).all(), ext
    finally:
        ref.refh_use_mi(b._handle, None, None)
        gcs.gamera_cs_mi_dipole_destroy(h)
        gcs.gcs_dropin_release(ref.refh_grid(b._handle), ref.refh_storage(b._handle))


def test_point_queries_match_the_reference_on_one_potential(libs):
    """gamera_cs_mi_dipole_potential, _ghost_velocity and _apply_inner_emf of
    a libgcs.so potential view holding the reference's solved potential
    against the reference's functions at random points and on a wall."""
    ref, gcs = libs
    cfg = mi_config(gcs_config.workload("small_mi"))
    a = problem.Problem(ref, "refh", cfg)
    a.advance(30)  # an update at t = 0 with the developing current
    ref_mi = ref.refh_mi(a._handle)
    pot, epoch, nxt = export_potential(ref, ref_mi, cfg)
    view = VP()
    assert gcs.gamera_cs_mi_dipole_create_potential_view(ref.refh_startup(a._handle),
                                                          ctypes.byref(cfg.mi),
                                                          ctypes.byref(view)) == 0
    try:
        assert gcs.gamera_cs_mi_dipole_ready(view) == 0
        assert gcs.gamera_cs_mi_dipole_import_potential(view, pot.ctypes.data, pot.size,
                                                        epoch, max(nxt, epoch + 1.0)) == 0
        assert gcs.gamera_cs_mi_dipole_ready(view) == 1
        rng = np.random.default_rng(5)
        worst_p = worst_v = 0.0
        for _ in range(200):
            d = rng.normal(size=3)
            x = Vec3((ctypes.c_double * 3)(*(d / np.linalg.norm(d) * rng.uniform(1.0, 3.0))))
            v = Vec3((ctypes.c_double * 3)(*rng.normal(size=3)))
            pa, pb = ctypes.c_double(), ctypes.c_double()
            ra = ref.gamera_cs_mi_dipole_potential(ref_mi, x, ctypes.byref(pa))
            rb = gcs.gamera_cs_mi_dipole_potential(view, x, ctypes.byref(pb))
            assert ra == rb
            worst_p = max(worst_p, abs(pa.value - pb.value))
            ga, gb = Vec3(), Vec3()
            ra = ref.gamera_cs_mi_dipole_ghost_velocity(ref_mi, x, v, ctypes.byref(ga))
            rb = gcs.gamera_cs_mi_dipole_ghost_velocity(view, x, v, ctypes.byref(gb))
            assert ra == rb
            if ra == 0:
                worst_v = max(worst_v, float(np.abs(np.array(ga.value) - np.array(gb.value)).max()
                                             / max(1.0, np.abs(np.array(ga.value)).max())))
        scale = np.abs(pot).max()
        print(f"point queries: potential |diff| {worst_p:.2e} V of {scale:.2e}; ghost velocity "
              f"scaled {worst_v:.2e}", flush=True)
        assert worst_p <= 1e-12 * scale
        assert worst_v <= 1e-12
        # the wall EMF of face 0's storage at the inner wall shell
        storage = ref.refh_storage(a._handle)
        grid = ref.refh_grid(a._handle)
        walls = []
        for lib, mi in ((ref, ref_mi), (gcs, view)):
            for d in (1, 2):
                a.field(0, problem.EDGE_EMF + d)[:] = 0.0
            assert lib.gamera_cs_mi_dipole_apply_inner_emf(mi, storage, grid, cfg.ghost) == 0
            walls.append([a.field(0, problem.EDGE_EMF + d).copy() for d in (1, 2)])
        for w, g in zip(*walls):
            assert (w != 0.0).any() and np.abs(w - g).max() <= 1e-12 * np.abs(w).max()
    finally:
        gcs.gamera_cs_mi_dipole_destroy(view)


def test_resident_driver_loop_moves_no_state(libs):
    """The reference's driver loop with libgcs.so's step, ionosphere and
    local CFL dt in resident mode: 20 steps with no host round trip, every
    dt after the first served from the device state, and the state synced
    once at the end equal to the reference's."""
    ref, gcs = libs
    cfg = mi_config(gcs_config.workload("small_mi"))
    a = problem.Problem(ref, "refh", cfg)
    b = problem.Problem(ref, "refh", cfg)
    ref.refh_set_step(b._handle, ctypes.cast(gcs.gamera_cs_advance_six_with_background_planned,
                                             VP).value)
    ref.refh_set_timestep.argtypes = [VP, VP]
    ref.refh_set_timestep(b._handle, ctypes.cast(gcs.gamera_no_local_timestep, VP).value)
    h = create_gcs_mi(ref, gcs, b, cfg)
    ref.refh_use_mi(b._handle, h, ctypes.cast(gcs.gamera_cs_mi_dipole_maybe_update, VP).value)
    gcs.gcs_dropin_local_timestep_device_calls.restype = ctypes.c_long
    gcs.gcs_dropin_sync_to_host.argtypes = [VP, VP]
    rounds = gcs.gcs_dropin_host_round_trips()
    served = gcs.gcs_dropin_local_timestep_device_calls()
    steps = 20
    assert gcs.gcs_dropin_set_resident(1) == 0
    try:
        for s in range(steps):
            a.advance(1)
            b.advance(1)
            ca, cb = a.clock(), b.clock()
            assert abs(ca["dt"] - cb["dt"]) <= STEP_TOL * (s + 1) * ca["dt"], s
        assert gcs.gcs_dropin_sync_to_host(ref.refh_grid(b._handle),
                                           ref.refh_storage(b._handle)) == 0
        worst = max(scaled_error(a.field(f, fid), b.field(f, fid))
                    for fid in problem.STATE_FIELDS for f in range(gcs_config.FACE_COUNT))
        rel = max(state_relative_errors(a, b).values())
        served = gcs.gcs_dropin_local_timestep_device_calls() - served
        print(f"resident loop: {steps} steps scaled {worst:.2e} relative {rel:.2e}; "
              f"dt calls served on the device {served}; host round trips "
              f"{gcs.gcs_dropin_host_round_trips() - rounds}", flush=True)
        assert worst <= STEP_TOL * steps
        assert served >= 6 * (steps - 1)
        assert gcs.gcs_dropin_host_round_trips() == rounds
    finally:
        gcs.gcs_dropin_set_resident(0)
        ref.refh_use_mi(b._handle, None, None)
        gcs.gamera_cs_mi_dipole_destroy(h)
        gcs.gcs_dropin_release(ref.refh_grid(b._handle), ref.refh_storage(b._handle))
