"""Shared fixtures: the package, the oracle and the compiled reference.

Markers: `gpu` tests need an MI355X and the built libgcs.so; everything else
runs on the CPU.
"""
import ctypes
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

gcs_config = importlib.import_module("gamera-cube_amd.config")
problem = importlib.import_module("gamera-cube_amd.problem")

ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libgamera_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def orc():
    if not os.path.exists(ORACLE_LIB):
        pytest.fail(f"{ORACLE_LIB} missing: run `make -C oracle`")
    return ctypes.CDLL(ORACLE_LIB)


@pytest.fixture(scope="session")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built here")
    return ctypes.CDLL(REF_LIB)


@pytest.fixture(scope="session")
def gpu():
    native = importlib.import_module("gamera-cube_amd.native")
    native.library()
    return native


def scaled_error(a, b):
    """max |a - b| / max(1, |a|): the reference's parity measure
    (cubedsphere/tests/check_tiled_mi_earth_parity.py:96-108)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(a))))


def relative_error(a, b, ncomp=1):
    """max over the components c of max |a_c - b_c| / max |a_c|: the error
    relative to each variable's own largest magnitude, which stays relative
    for arrays whose values are all below 1 (residual B, face fluxes, EMFs),
    where the scaled measure above is absolute."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, ncomp)
    b = np.asarray(b, dtype=np.float64).reshape(-1, ncomp)
    if a.size == 0:
        return 0.0
    scale = np.abs(a).max(axis=0)
    diff = np.abs(a - b).max(axis=0)
    return float(np.max(np.where(scale > 0.0, diff / np.where(scale > 0.0, scale, 1.0),
                                 diff)))


# components per element of the state fields (problem.STATE_FIELDS order)
STATE_COMPONENTS = [5, 5, 3, 3, 1, 1, 1, 1, 1, 1]

# The state's physical quantities: (field ids, components per element,
# component slices), each quantity's scale being its largest magnitude over
# all six faces (and, for the face flux, all three directions).
STATE_QUANTITIES = {
    "rho": ([problem.CONSERVED], 5, slice(0, 1)),
    "momentum": ([problem.CONSERVED], 5, slice(1, 4)),
    "energy": ([problem.CONSERVED], 5, slice(4, 5)),
    "old rho": ([problem.OLD_CONSERVED], 5, slice(0, 1)),
    "old momentum": ([problem.OLD_CONSERVED], 5, slice(1, 4)),
    "old energy": ([problem.OLD_CONSERVED], 5, slice(4, 5)),
    "B": ([problem.CELL_MAGNETIC], 3, slice(0, 3)),
    "old B": ([problem.OLD_CELL_MAGNETIC], 3, slice(0, 3)),
    "face flux": ([problem.FACE_FLUX + d for d in range(3)], 1, slice(0, 1)),
    "old face flux": ([problem.OLD_FACE_FLUX + d for d in range(3)], 1, slice(0, 1)),
}


def state_relative_errors(a, b):
    """{quantity: max |a - b| / max |a|} over the state of two problems:
    the error of each conserved variable (density, momentum vector, energy),
    the residual B vector and the face flux relative to that quantity's own
    scale. Unlike the scaled measure, it stays relative for quantities whose
    values are all below 1; unlike a per-component one, it does not divide
    by a component that is zero up to rounding (a momentum component
    across the flow at the start of a run)."""
    out = {}
    for name, (fids, nc, comp) in STATE_QUANTITIES.items():
        scale = diff = 0.0
        for fid in fids:
            for f in range(gcs_config.FACE_COUNT):
                x = np.asarray(a.field(f, fid)).reshape(-1, nc)[:, comp]
                y = np.asarray(b.field(f, fid)).reshape(-1, nc)[:, comp]
                scale = max(scale, float(np.abs(x).max()))
                diff = max(diff, float(np.abs(x - y).max()))
        out[name] = diff / scale if scale > 0.0 else diff
    return out


def counters(p, face):
    """(emergency-diffusion face count, maximum interface speed) that the
    steps so far accumulated into one cube face's storage
    (nonorthogonal_storage.h:30-31), from any of the three problem APIs."""
    if hasattr(p, "counters"):
        return tuple(p.counters(face))
    out = (ctypes.c_double * 2)()
    fn = getattr(p._lib, f"{p.prefix}_counters")
    fn.restype = None
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    fn(p._handle, face, out)
    return (out[0], out[1])


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def oracle_from(source, orc_lib, cfg):
    """An oracle problem holding a copy of `source`'s inputs."""
    o = problem.Problem(orc_lib, "orc", cfg)
    o.copy_fields(source, problem.GEOMETRY_FIELDS + problem.BACKGROUND_FIELDS
                  + problem.STATE_FIELDS)
    for f in range(gcs_config.FACE_COUNT):
        for d in range(3):
            o.edge_valid(f, d)[:] = source.edge_valid(f, d)
    return o


def reference_from(source, ref_lib, cfg):
    """The compiled reference's problem holding a copy of `source`'s inputs
    (allocated by refh_create_shell, so the reference's setup is skipped);
    its clock continues `source`'s."""
    r = problem.Problem(ref_lib, "refh", cfg, create="create_shell")
    r.copy_fields(source, problem.GEOMETRY_FIELDS + problem.BACKGROUND_FIELDS
                  + problem.STATE_FIELDS)
    for f in range(gcs_config.FACE_COUNT):
        for d in range(3):
            r.edge_valid(f, d)[:] = source.edge_valid(f, d)
    c = source.clock()
    fn = ref_lib.refh_set_clock
    fn.restype = None
    fn.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.c_long]
    fn(r._handle, c["time"], c["dt"], c["previous_dt"], c["steps"])
    return r


def active_view(p, face, field_id, cfg, ncomp):
    """Active cells of a cell field, shape (nr, n, n, ncomp)."""
    ni, nj, nk = cfg.extent
    g = cfg.ghost
    a = p.field(face, field_id).reshape(ni, nj, nk, ncomp)
    return a[g:g + cfg.radial_cells, g:g + cfg.angular_cells,
             g:g + cfg.angular_cells]


def max_divergence(p, cfg):
    """max over active cells of |net total (B1 + B0) face flux| / volume:
    the reference's gate (test_cubedsphere_dipole_background.c:808-845)."""
    ni, nj, nk = cfg.extent
    g, nr, n = cfg.ghost, cfg.radial_cells, cfg.angular_cells
    worst = 0.0
    for f in range(gcs_config.FACE_COUNT):
        vol = p.field(f, problem.CELL_GEOMETRY).reshape(ni, nj, nk, 103)[..., 0]
        net = np.zeros((nr, n, n))
        for d in range(3):
            ext = [ni, nj, nk]
            ext[d] += 1
            phi = (p.field(f, problem.FACE_FLUX + d)
                   + p.field(f, problem.BG_FACE_FLUX + d)).reshape(ext)
            sl = [slice(g, g + nr), slice(g, g + n), slice(g, g + n)]
            hi = list(sl)
            hi[d] = slice(sl[d].start + 1, sl[d].stop + 1)
            net += phi[tuple(hi)] - phi[tuple(sl)]
        worst = max(worst, float(np.max(np.abs(net) / vol[g:g + nr, g:g + n, g:g + n])))
    return worst


def wind_series(samples=40, seed=11):
    """A long solar-wind series for the step's Earth boundary: times from
    -20 to about +3.4 code units (the ballistic delay across the grid spans
    about 15), density, pressure, speed and IMF varying between samples,
    Bz switching sign in blocks. Returns (time, state[n][8] = rho, p, v, B)."""
    rng = np.random.default_rng(seed)
    time = -20.0 + 0.6 * np.arange(samples)
    state = np.zeros((samples, 8))
    state[:, 0] = 5.0 + rng.uniform(-2.0, 2.0, samples)
    state[:, 1] = state[:, 0] * 0.08 * (1.0 + rng.uniform(-0.2, 0.2, samples))
    state[:, 2] = 4.0 + rng.uniform(-0.5, 0.5, samples)
    state[:, 5] = rng.uniform(-0.1, 0.1, samples)
    state[:, 6] = rng.uniform(-0.3, 0.3, samples)
    state[:, 7] = np.where((np.arange(samples) // 5) % 2 == 0, -1.09, 1.09)
    return time, state


def set_wind(p, time, state, linear=True):
    """Give a problem of any of the three APIs the wind series (time, state)."""
    fn = getattr(p._lib, f"{p.prefix}_set_wind")
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_int]
    t = np.ascontiguousarray(time, dtype=np.float64)
    s = np.ascontiguousarray(state, dtype=np.float64)
    rc = fn(p._handle, t.size, t.ctypes.data, s.ctypes.data, 1 if linear else 0)
    if rc != 0:
        raise RuntimeError(f"{p.prefix}: set_wind failed")
