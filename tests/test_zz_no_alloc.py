"""gcs_problem_pressure_control_noalloc (static device scratch, no device
allocation) against gcs_problem_pressure_control on identical problems:
counts, maxima and the controlled state bit for bit, on two calls in a row,
the second on a smaller problem, and the size limit's error."""
import ctypes

import pytest

from conftest import bits_equal, gcs_config, problem
from test_pressure_control import GAMMA, call, state_parameters

pytestmark = pytest.mark.gpu


def controlled_pair(gpu, name):
    """Two identical problems after three steps, one controlled by each
    entry point with thresholds that make both branches fire."""
    cfg = gcs_config.workload(name)
    a, b = gpu.GpuProblem(cfg), gpu.GpuProblem(cfg)
    a.advance(3)
    b.advance(3)
    params = state_parameters(a, cfg, GAMMA)
    dt = a.clock()["dt"]
    lib = gpu.library()
    got_a = call(lib, "gcs_problem", a._handle, dt, GAMMA, params)
    fn = lib.gcs_problem_pressure_control_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = ([ctypes.c_void_p] + [ctypes.c_double] * 7
                   + [ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_double)])
    counts = (ctypes.c_ulonglong * 2)()
    maxima = (ctypes.c_double * 2)()
    rc = fn(b._handle, dt, GAMMA, params["density_floor"], params["pressure_floor"],
            params["light_speed"], params["low_density_threshold"],
            params["low_density_sound_speed"], counts, maxima)
    got_b = (rc, tuple(counts), tuple(maxima))
    return cfg, a, b, got_a, got_b


@pytest.mark.parametrize("name", ["small", "tiny"])  # the second call is smaller
def test_noalloc_pressure_control_matches_allocating_entry_point(gpu, name):
    cfg, a, b, got_a, got_b = controlled_pair(gpu, name)
    print(f"{name}: allocating {got_a}, no-alloc {got_b}", flush=True)
    assert got_a[0] == got_b[0] == 0
    assert got_a[1] == got_b[1] and got_a[1][0] > 0 and got_a[1][1] > 0
    assert bits_equal(got_a[2], got_b[2])
    for f in range(gcs_config.FACE_COUNT):
        assert bits_equal(a.field(f, problem.CONSERVED), b.field(f, problem.CONSERVED)), f
    # both continue identically
    a.advance(1)
    b.advance(1)
    for f in range(gcs_config.FACE_COUNT):
        assert bits_equal(a.field(f, problem.CONSERVED), b.field(f, problem.CONSERVED)), f


def test_noalloc_pressure_control_rejects_problems_above_its_scratch(gpu):
    """6 x 116^2 x 208 = 16,792,704 active cells: one block of 256 more than
    the static scratch holds."""
    cfg = gcs_config.earth_problem(208, 116)
    g = gpu.GpuProblem(cfg)
    fn = gpu.library().gcs_problem_pressure_control_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = ([ctypes.c_void_p] + [ctypes.c_double] * 7
                   + [ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_double)])
    counts = (ctypes.c_ulonglong * 2)()
    maxima = (ctypes.c_double * 2)()
    rc = fn(g._handle, 1e-3, GAMMA, 1e-6, 1e-8, 10.0, 1e-3, 1e-7, counts, maxima)
    assert rc == -1
    assert "scratch limit" in gpu.last_error()
