"""The drop-in entry points (include/gamera_cs_dropin.h) called by the
compiled reference's own driver loop in place of its step.

Two copies of one reference problem run side by side: one steps with the
reference's gamera_cs_advance_six_with_background_planned, the other with
libgcs.so's function of the same prototype (the harness calls it through a
pointer; the reference's dt rule, Earth context and storage are used
unchanged). Every array the step writes is compared after each step with the
reference's scaled measure |a - b| / max(1, |a|).

The stress rates (hydro_rate, maxwell_rate) of the active cells, which the
device update stores when it reproduces the storage, are compared as well.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import REF_LIB, gcs_config, problem, scaled_error

pytestmark = pytest.mark.gpu

# Per step, as for the problem API (ulp differences of ROCm's transcendental
# functions from glibc's grow by at most about 1e-14 per step).
STEP_TOL = 1e-13

WRITTEN = ([problem.CONSERVED, problem.OLD_CONSERVED, problem.PREDICTED_CONSERVED,
            problem.CELL_MAGNETIC, problem.OLD_CELL_MAGNETIC,
            problem.PREDICTED_CELL_MAGNETIC]
           + [problem.FACE_FLUX + d for d in range(3)]
           + [problem.OLD_FACE_FLUX + d for d in range(3)]
           + [problem.PREDICTED_FACE_FLUX + d for d in range(3)]
           + [problem.EDGE_EMF + d for d in range(3)]
           + [problem.FLUID_FACE_FLUX + d for d in range(3)]
           + [problem.MAXWELL_FACE_FLUX + d for d in range(3)]
           + [problem.HYDRO_RATE, problem.MAXWELL_RATE])


@pytest.fixture(scope="module")
def ref_global():
    """The compiled reference, loaded into the global symbol table so that
    libgcs.so finds its Earth callbacks by name."""
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    lib = ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_GLOBAL)
    lib.refh_set_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.refh_use_wrapped_callbacks.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.refh_step_plain.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_double, ctypes.c_double]
    lib.refh_grid.restype = ctypes.c_void_p
    lib.refh_grid.argtypes = [ctypes.c_void_p]
    lib.refh_storage.restype = ctypes.c_void_p
    lib.refh_storage.argtypes = [ctypes.c_void_p]
    return lib


def address(lib, name):
    return ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


# Area-scaled face fluxes are intermediates whose terms cancel: an interface
# flux is the sum of two half-fluxes of size area * U * (|v| + c_s), and its
# pressure terms come from p = (gamma - 1) (E - rho v^2 / 2). In a cold
# hypersonic stream (the 1600 km/s wind: E ~ 640, p ~ 0.4, an angular
# energy flux of 0.1) a state difference at the step tolerance moves them by
# far more than their own values' roundoff. They are measured against the
# largest flux of the block: every component and face direction of the
# same array kind (the stream-wise energy flux carries area * E * v). The
# stress rates are measured against the array's largest value.
FACE_FLUXES = (problem.FLUID_FACE_FLUX, problem.MAXWELL_FACE_FLUX)
INTERMEDIATE = ([base + d for base in FACE_FLUXES for d in range(3)]
                + [problem.HYDRO_RATE, problem.MAXWELL_RATE])


def flux_scale(p, face, fid):
    base = problem.FLUID_FACE_FLUX if fid < problem.MAXWELL_FACE_FLUX else \
        problem.MAXWELL_FACE_FLUX
    return max(float(np.abs(p.field(face, base + d)).max(initial=0.0)) for d in range(3))


def worst(a, b, fields=WRITTEN):
    out = (0.0, None)
    for fid in fields:
        for f in range(gcs_config.FACE_COUNT):
            x, y = a.field(f, fid), b.field(f, fid)
            if not x.size:
                continue
            if fid in (problem.HYDRO_RATE, problem.MAXWELL_RATE):
                e = float(np.abs(x - y).max() / max(1.0, np.abs(x).max()))
            elif fid in INTERMEDIATE:
                e = float(np.abs(x - y).max() / max(1.0, flux_scale(a, f, fid)))
            else:
                e = scaled_error(x, y)
            if e > out[0]:
                out = (e, (fid, f))
    return out


def release(gpu_lib, ref_lib, p):
    gpu_lib.gcs_dropin_release.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    gpu_lib.gcs_dropin_release(ref_lib.refh_grid(p._handle),
                               ref_lib.refh_storage(p._handle))


def run_pair(ref_lib, gpu, name, steps, wrapped, step=None, mhd=True, tile_emf=False):
    """`step`: address of a function with the planned entry point's
    prototype (default: libgcs.so's own); mhd=False: hydrodynamic steps;
    tile_emf: the MPI Earth driver's EMF callback on one radial tile.
    Returns the host round trips the drop-in made."""
    lib = gpu.library()
    lib.gcs_dropin_host_round_trips.restype = ctypes.c_long
    rounds = lib.gcs_dropin_host_round_trips()
    cfg = gcs_config.workload(name)
    a = problem.Problem(ref_lib, "refh", cfg)
    b = problem.Problem(ref_lib, "refh", cfg)
    if not mhd:
        ref_lib.refh_set_mhd.argtypes = [ctypes.c_void_p, ctypes.c_int]
        ref_lib.refh_set_mhd(a._handle, 0)
        ref_lib.refh_set_mhd(b._handle, 0)
    if tile_emf:
        ref_lib.refh_use_tile_emf.argtypes = [ctypes.c_void_p, ctypes.c_int]
        ref_lib.refh_use_tile_emf(a._handle, 1)
        ref_lib.refh_use_tile_emf(b._handle, 1)
    if step is None:
        step = address(lib, "gamera_cs_advance_six_with_background_planned")
    ref_lib.refh_set_step(b._handle, step)
    if wrapped:
        ref_lib.refh_use_wrapped_callbacks(a._handle, 1)
        ref_lib.refh_use_wrapped_callbacks(b._handle, 1)
    try:
        for s in range(steps):
            a.advance(1)
            try:
                b.advance(1)
            except RuntimeError:
                pytest.fail(f"drop-in step {s + 1} failed: {gpu.last_error()}")
            err, where = worst(a, b)
            ca, cb = a.clock(), b.clock()
            print(f"dropin {name} wrapped={wrapped} step {s + 1} err {err:.3e} at {where}",
                  flush=True)
            assert err <= STEP_TOL * (s + 1), (s + 1, where)
            assert abs(ca["dt"] - cb["dt"]) <= STEP_TOL * (s + 1) * ca["dt"]
    finally:
        release(lib, ref_lib, b)
    return lib.gcs_dropin_host_round_trips() - rounds


@pytest.mark.parametrize("name,steps", [("tiny", 4), ("small", 8), ("ghost5", 4)])
def test_dropin_matches_reference_step(ref_global, gpu, name, steps):
    """Earth callbacks recognised by address and run as device kernels
    (ghost5: five ghost layers, which the reference allows)."""
    run_pair(ref_global, gpu, name, steps, wrapped=False)


@pytest.mark.parametrize("name,steps", [("small", 6), ("fast_wind", 4)])
def test_dropin_hydrodynamic_steps_match_reference(ref_global, gpu, name, steps):
    """use_mhd = false (nonorthogonal_step.c:554, nonorthogonal_sweep.c:
    150-171): fluid fluxes only, no EMF or CT; B and the face fluxes are
    carried over unchanged."""
    run_pair(ref_global, gpu, name, steps, wrapped=False, mhd=False)


# The planned entry point's prototype (cubedsphere_step.h:58-67).
PLANNED = ctypes.CFUNCTYPE(
    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
    ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p)

MPI_ENTRY_POINTS = [
    "gamera_cs_mpi_advance_six_with_background_planned",
    "gamera_cs_mpi_advance_six_with_background_planned_cached_edges",
    "gamera_cs_mpi_advance_six_with_background_planned_cached_edges_pre_synced",
    "gamera_cs_mpi_advance_six_with_background_sparse_pre_synced",
]


def mpi_as_planned(lib, name):
    """The MPI entry point `name` (include/gamera_cs_mpi_dropin.h) behind
    the planned prototype, with no edge plan and the one-process shim
    communicator."""
    fn = getattr(lib, name)
    edge_plan = name != MPI_ENTRY_POINTS[0]
    v = ctypes.c_void_p
    fn.restype = ctypes.c_int
    fn.argtypes = ([v, v, v, ctypes.c_size_t, ctypes.c_double, ctypes.c_double, v, v, v]
                   + ([v] if edge_plan else [])
                   + [v, ctypes.c_double, ctypes.c_double, v, v])

    def step(grid, storage, active, ghost, ratio, dt, options, background, plan, radial,
             t, t_prev, context):
        extra = [None] if edge_plan else []
        return fn(grid, storage, active, ghost, ratio, dt, options, background, plan,
                  *extra, radial, t, t_prev, context, None)
    return PLANNED(step)


@pytest.mark.parametrize("entry", MPI_ENTRY_POINTS)
def test_mpi_entry_points_match_reference_step(ref_global, gpu, entry):
    """cubedsphere_mpi_step.h:36-91 with a one-process communicator: the
    serial step's results (the reference gates its MPI step on exact
    agreement with the serial one)."""
    adapter = mpi_as_planned(gpu.library(), entry)
    run_pair(ref_global, gpu, "small", 3, wrapped=False,
             step=ctypes.cast(adapter, ctypes.c_void_p).value)


def test_dropin_runs_unknown_callbacks_on_host(ref_global, gpu):
    """Callbacks at other addresses run on the host between device phases."""
    assert run_pair(ref_global, gpu, "small", 3, wrapped=True) > 0


def test_tile_emf_callback_runs_on_device(ref_global, gpu):
    """gamera_cs_earth_tile_emf (cubedsphere_earth_boundary.c:453-464), the
    MPI Earth driver's EMF callback, on one radial tile applies both wall
    EMFs: the device runs it like gamera_cs_earth_no_mi_emf, with the tile
    context's Earth context, and no state goes to the host."""
    assert run_pair(ref_global, gpu, "small", 4, wrapped=False, tile_emf=True) == 0


def test_plain_entry_point_matches_reference(ref_global, gpu):
    """gamera_cs_advance_six: no background, radial outflow ghosts."""
    lib = gpu.library()
    cfg = gcs_config.workload("small")
    a = problem.Problem(ref_global, "refh", cfg)
    b = problem.Problem(ref_global, "refh", cfg)
    plain = address(lib, "gamera_cs_advance_six")
    dt = a.clock()["dt"]
    try:
        for s, ratio in enumerate((0.0, 0.5, 0.5, 0.5)):
            assert ref_global.refh_step_plain(a._handle, None, ratio, dt) == 0
            if ref_global.refh_step_plain(b._handle, plain, ratio, dt) != 0:
                pytest.fail(f"plain drop-in step failed: {gpu.last_error()}")
            err, where = worst(a, b)
            print(f"plain step {s + 1} err {err:.3e} at {where}", flush=True)
            assert err <= STEP_TOL * (s + 1), where
    finally:
        release(lib, ref_global, b)
    assert lib.gcs_dropin_mirror_count() == 0


def test_invalid_arguments_are_rejected_before_writing(ref_global, gpu):
    """cubedsphere_step.c:23-47: a non-finite dt returns -1 and leaves the
    storage untouched."""
    lib = gpu.library()
    cfg = gcs_config.workload("tiny")
    b = problem.Problem(ref_global, "refh", cfg)
    before = [b.field(f, problem.CONSERVED).copy() for f in range(6)]
    plain = address(lib, "gamera_cs_advance_six")
    assert ref_global.refh_step_plain(b._handle, plain, 0.0, float("nan")) == -1
    assert ref_global.refh_step_plain(b._handle, plain, float("inf"), 0.1) == -1
    for f in range(6):
        assert (b.field(f, problem.CONSERVED) == before[f]).all()
    assert lib.gcs_dropin_mirror_count() == 0


class TimestepOptions(ctypes.Structure):
    """gamera_no_timestep_options (nonorthogonal_step.h:66-76)."""
    _fields_ = [("gamma", ctypes.c_double), ("density_floor", ctypes.c_double),
                ("pressure_floor", ctypes.c_double), ("cfl", ctypes.c_double),
                ("light_speed", ctypes.c_double), ("use_mhd", ctypes.c_bool),
                ("use_boris", ctypes.c_bool), ("use_background", ctypes.c_bool)]


def test_local_timestep_matches_reference(ref_global, gpu):
    """gamera_no_local_timestep (nonorthogonal_step.c:449-525) of libgcs.so
    against the reference's on every face of a stepped problem, for MHD and
    hydrodynamic options, with and without Boris and background, on the
    active box and a ragged sub-box; invalid options, boxes and states give
    -1 from both."""
    lib = gpu.library()
    cfg = gcs_config.workload("small")
    a = problem.Problem(ref_global, "refh", cfg)
    a.advance(3)
    fn = ref_global.refh_local_timestep
    fn.restype = ctypes.c_int
    size3 = ctypes.c_size_t * 3
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, size3, size3,
                   ctypes.POINTER(TimestepOptions), ctypes.POINTER(ctypes.c_double)]
    ours = address(lib, "gamera_no_local_timestep")
    g = cfg.ghost
    active = (size3(g, g, g), size3(g + cfg.radial_cells, g + cfg.angular_cells,
                                    g + cfg.angular_cells))
    ragged = (size3(g + 1, g + 3, g), size3(g + cfg.radial_cells - 2, g + 5,
                                            g + cfg.angular_cells - 1))
    extent = [cfg.radial_cells + 2 * g, cfg.angular_cells + 2 * g, cfg.angular_cells + 2 * g]

    def both(face, box, opts):
        da, db = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
        ra = fn(a._handle, None, face, box[0], box[1], ctypes.byref(opts), ctypes.byref(da))
        rb = fn(a._handle, ours, face, box[0], box[1], ctypes.byref(opts), ctypes.byref(db))
        return ra, rb, da.value, db.value

    worst_rel = 0.0
    for mhd, boris, bg in [(True, True, True), (True, False, True), (True, True, False),
                           (False, False, False)]:
        opts = TimestepOptions(5.0 / 3.0, 1e-6, 1e-8, cfg.cfl, 10.0, mhd, boris, bg)
        for face in range(6):
            for box in (active, ragged):
                ra, rb, da, db = both(face, box, opts)
                assert ra == rb == 0, (mhd, boris, bg, face)
                worst_rel = max(worst_rel, abs(da - db) / da)
    print(f"local timestep: worst relative difference {worst_rel:.2e}", flush=True)
    assert worst_rel <= 1e-14
    # nonorthogonal_step.c:454-471: options and boxes rejected by both
    bad = TimestepOptions(5.0 / 3.0, 1e-6, 1e-8, 0.0, 10.0, True, True, True)
    assert both(0, active, bad)[:2] == (-1, -1)
    empty = (size3(g, g, g), size3(g, g + 4, g + 4))
    good = TimestepOptions(5.0 / 3.0, 1e-6, 1e-8, cfg.cfl, 10.0, True, True, True)
    assert both(0, empty, good)[:2] == (-1, -1)
    beyond = (size3(g, g, g), size3(extent[0] + 1, extent[1], extent[2]))
    assert both(0, beyond, good)[:2] == (-1, -1)
    # an invalid state inside the box
    u = a.field(5, problem.CONSERVED).reshape(-1, 5)
    u[(g * extent[1] + g) * extent[2] + g, 0] = -1.0
    assert both(5, active, good)[:2] == (-1, -1)
    assert lib.gcs_dropin_mirror_count() == 0
