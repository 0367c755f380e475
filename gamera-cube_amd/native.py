"""Loader for the in-tree HIP library (gamera-cube_amd/libgcs.so).

There is no CPU fallback: if the library is missing or the GPU cannot run it,
these calls raise.
"""
import ctypes
import os

from .config import ProblemConfig  # noqa: F401
from .problem import Problem

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgcs.so")

_lib = None


def library():
    """Load libgcs.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: run `python gamera-cube_amd/build.py`")
        lib = ctypes.CDLL(LIB_PATH)
        lib.gcs_last_error.restype = ctypes.c_char_p
        lib.gcs_config_size.restype = ctypes.c_size_t
        if lib.gcs_config_size() != ctypes.sizeof(ProblemConfig):
            raise RuntimeError("ProblemConfig does not match gcs_problem_config")
        lib.gcs_problem_counters.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double)]
        _lib = lib
    return _lib


def last_error():
    return library().gcs_last_error().decode()


class GpuProblem(Problem):
    """A six-face Earth problem resident on the GPU."""

    def __init__(self, config):
        lib = library()
        try:
            super().__init__(lib, "gcs_problem", config)
        except RuntimeError as exc:
            raise RuntimeError(f"{exc}: {last_error()}") from None

    def advance(self, steps):
        try:
            super().advance(steps)
        except RuntimeError as exc:
            raise RuntimeError(f"{exc}: {last_error()}") from None

    def step_explicit(self, ratio, dt, time, previous_time):
        try:
            super().step_explicit(ratio, dt, time, previous_time)
        except RuntimeError as exc:
            raise RuntimeError(f"{exc}: {last_error()}") from None

    def counters(self, face):
        out = (ctypes.c_double * 2)()
        self._lib.gcs_problem_counters(self._handle, face, out)
        return out[0], out[1]


def stage_names():
    lib = library()
    lib.gcs_stage_name.restype = ctypes.c_char_p
    return [lib.gcs_stage_name(i).decode() for i in range(lib.gcs_stage_count())]


def _time_steps(self, steps, profile=True):
    """(total ms, {stage: ms summed over the steps}) timed with HIP events."""
    lib = self._lib
    n = lib.gcs_stage_count()
    elapsed = ctypes.c_double()
    stages = (ctypes.c_double * n)()
    fn = lib.gcs_problem_time_steps
    fn.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    if fn(self._handle, int(steps), 1 if profile else 0, ctypes.byref(elapsed),
          stages) != 0:
        raise RuntimeError(f"gcs_problem_time_steps failed: {last_error()}")
    names = stage_names()
    return elapsed.value, {names[i]: stages[i] for i in range(n)} if profile else {}


def _set_fidelity(self, on):
    self._lib.gcs_problem_set_fidelity.argtypes = [ctypes.c_void_p, ctypes.c_int]
    if self._lib.gcs_problem_set_fidelity(self._handle, 1 if on else 0) != 0:
        raise RuntimeError("gcs_problem_set_fidelity failed")


GpuProblem.time_steps = _time_steps
GpuProblem.set_fidelity = _set_fidelity


def _mi_report(self):
    """Ionosphere coupling: updates, readiness, iterations and residuals of
    the last solve per hemisphere, its epoch (s) and duration (ms)."""
    lib = self._lib
    ints = (ctypes.c_int * 4)()
    dbl = (ctypes.c_double * 4)()
    lib.gcs_problem_mi_report.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    if lib.gcs_problem_mi_report(self._handle, ints, dbl) != 0:
        raise RuntimeError(f"gcs_problem_mi_report: {last_error()}")
    return {"updates": ints[0], "ready": bool(ints[1]),
            "iterations": (ints[2], ints[3]), "final_residual": (dbl[0], dbl[1]),
            "epoch_s": dbl[2], "solve_ms": dbl[3]}


def _mi_field(self, what):
    """Copy of the potential (what=0, V) or sampled FAC (what=1, A/m^2),
    [hemisphere][colatitude][azimuth]."""
    import numpy as np
    lib = self._lib
    ptr = ctypes.POINTER(ctypes.c_double)()
    count = ctypes.c_size_t()
    lib.gcs_problem_mi_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p]
    if lib.gcs_problem_mi_field(self._handle, what, ctypes.byref(ptr),
                                ctypes.byref(count)) != 0:
        raise RuntimeError(f"gcs_problem_mi_field: {last_error()}")
    return np.ctypeslib.as_array(ptr, shape=(count.value,)).copy()


def _mi_sample_current(self):
    """Sample the current state's field-aligned current onto the polar grids
    (no solve); mi_field(1) returns it."""
    self._lib.gcs_problem_mi_sample_current.argtypes = [ctypes.c_void_p]
    if self._lib.gcs_problem_mi_sample_current(self._handle) != 0:
        raise RuntimeError(f"gcs_problem_mi_sample_current: {last_error()}")


GpuProblem.mi_sample_current = _mi_sample_current
GpuProblem.mi_report = _mi_report
GpuProblem.mi_field = _mi_field


def _set_precision(self, mixed):
    """False: FP64 throughout. True: FP32 reconstruction/fluxes/EMF with
    FP32 intermediate arrays, FP64 update, CT, CFL and ionosphere."""
    self._lib.gcs_problem_set_precision.argtypes = [ctypes.c_void_p, ctypes.c_int]
    if self._lib.gcs_problem_set_precision(self._handle, 1 if mixed else 0) != 0:
        raise RuntimeError(f"gcs_problem_set_precision: {last_error()}")


GpuProblem.set_precision = _set_precision


def _pressure_control(self, dt, gamma, density_floor, pressure_floor, light_speed,
                      low_density_threshold, low_density_sound_speed):
    """gamera_no_apply_pressure_control on the current state (see
    include/gamera_cs_step.h); returns ((low-density cells, high-sound-speed
    cells), (maximum sound speed, maximum pressure))."""
    fn = self._lib.gcs_problem_pressure_control
    fn.restype = ctypes.c_int
    fn.argtypes = ([ctypes.c_void_p] + [ctypes.c_double] * 7
                   + [ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_double)])
    counts = (ctypes.c_ulonglong * 2)()
    maxima = (ctypes.c_double * 2)()
    if fn(self._handle, dt, gamma, density_floor, pressure_floor, light_speed,
          low_density_threshold, low_density_sound_speed, counts, maxima) != 0:
        raise RuntimeError(f"gcs_problem_pressure_control: {last_error()}")
    return (counts[0], counts[1]), (maxima[0], maxima[1])


GpuProblem.pressure_control = _pressure_control


def _set_graph(self, on):
    """True: each driver step is one hipGraph launch (same results)."""
    self._lib.gcs_problem_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    if self._lib.gcs_problem_set_graph(self._handle, 1 if on else 0) != 0:
        raise RuntimeError("gcs_problem_set_graph failed")


GpuProblem.set_graph = _set_graph


def _max_div_b(self):
    """max |div B| over the active cells (net total face flux / volume)."""
    out = ctypes.c_double()
    self._lib.gcs_problem_max_div_b.argtypes = [ctypes.c_void_p,
                                                ctypes.POINTER(ctypes.c_double)]
    if self._lib.gcs_problem_max_div_b(self._handle, ctypes.byref(out)) != 0:
        raise RuntimeError(f"gcs_problem_max_div_b: {last_error()}")
    return out.value


GpuProblem.max_div_b = _max_div_b


def _set_plain_kernels(self, on):
    """True: one face or edge per thread in the face-flux, face-velocity and
    EMF sweeps (no LDS tiles or register marches); results are identical."""
    self._lib.gcs_problem_set_plain_kernels.argtypes = [ctypes.c_void_p, ctypes.c_int]
    if self._lib.gcs_problem_set_plain_kernels(self._handle, 1 if on else 0) != 0:
        raise RuntimeError("gcs_problem_set_plain_kernels failed")


GpuProblem.set_plain_kernels = _set_plain_kernels
