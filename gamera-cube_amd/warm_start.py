"""Host side of the learned ionosphere warm start (include/gamera_cs_unet.h).

`WarmStartNet` owns a device U-Net; `attach`, `train`, `make_pairs` and
`solve_current` act on a GpuProblem's MI solve. Everything runs in
libgcs.so; this module only moves host arrays through the C ABI and raises
on any failure.
"""
import ctypes

import numpy as np

from .native import GpuProblem, last_error, library

IN_CHANNELS = 32

_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)


class TrainConfig(ctypes.Structure):
    """gcs_warm_train_config."""
    _fields_ = [("steps", ctypes.c_long), ("batch", ctypes.c_int),
                ("learning_rate", ctypes.c_double), ("seed", ctypes.c_ulonglong),
                ("pedersen_min", ctypes.c_double), ("pedersen_max", ctypes.c_double),
                ("hall_min", ctypes.c_double), ("hall_max", ctypes.c_double)]


def _declare(lib):
    if getattr(lib, "_unet_declared", False):
        return
    lib.gcs_unet_create.restype = ctypes.c_void_p
    lib.gcs_unet_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_ulonglong]
    lib.gcs_unet_destroy.argtypes = [ctypes.c_void_p]
    lib.gcs_unet_param_count.restype = ctypes.c_size_t
    lib.gcs_unet_param_count.argtypes = [ctypes.c_void_p]
    for name in ("gcs_unet_set_params", "gcs_unet_get_params", "gcs_unet_get_grads"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, _F32P, ctypes.c_size_t]
    lib.gcs_unet_output_scale.restype = ctypes.c_double
    lib.gcs_unet_output_scale.argtypes = [ctypes.c_void_p]
    lib.gcs_unet_set_output_scale.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.gcs_unet_forward.argtypes = [ctypes.c_void_p, _F32P, ctypes.c_int, _F32P]
    lib.gcs_unet_gradients.argtypes = [ctypes.c_void_p, _F32P, _F32P, ctypes.c_int,
                                       _F64P]
    lib.gcs_unet_adam_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.gcs_mfma_gemm_check.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, _F32P]
    lib.gcs_problem_mi_attach_warm_start.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.gcs_problem_mi_train_warm_start.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.POINTER(TrainConfig), _F32P]
    lib.gcs_problem_mi_train_warm_start_on.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(TrainConfig), _F64P, _F64P,
        ctypes.POINTER(ctypes.c_int), ctypes.c_int, _F32P]
    lib.gcs_problem_mi_make_pairs.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong,
                                              ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, _F64P, _F64P]
    lib.gcs_problem_mi_solve_current.argtypes = [ctypes.c_void_p, _F64P, _F64P,
                                                 ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_int), _F64P]
    lib._unet_declared = True


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {last_error()}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class WarmStartNet:
    """The 4-level U-Net on an H x W (colatitude x azimuth) grid."""

    def __init__(self, height, width, max_batch, seed=0):
        self._lib = library()
        _declare(self._lib)
        self.height, self.width, self.max_batch = height, width, max_batch
        self._handle = self._lib.gcs_unet_create(height, width, max_batch, seed)
        if not self._handle:
            raise RuntimeError(f"gcs_unet_create: {last_error()}")

    def close(self):
        if self._handle:
            self._lib.gcs_unet_destroy(self._handle)
            self._handle = None

    def __del__(self):
        self.close()

    @property
    def param_count(self):
        return self._lib.gcs_unet_param_count(self._handle)

    def params(self):
        out = np.zeros(self.param_count, dtype=np.float32)
        _check(self._lib.gcs_unet_get_params(self._handle, out.ctypes.data_as(_F32P),
                                             out.size), "gcs_unet_get_params")
        return out

    def set_params(self, flat):
        flat = _f32(flat)
        _check(self._lib.gcs_unet_set_params(self._handle, flat.ctypes.data_as(_F32P),
                                             flat.size), "gcs_unet_set_params")

    def grads(self):
        out = np.zeros(self.param_count, dtype=np.float32)
        _check(self._lib.gcs_unet_get_grads(self._handle, out.ctypes.data_as(_F32P),
                                            out.size), "gcs_unet_get_grads")
        return out

    @property
    def output_scale(self):
        return self._lib.gcs_unet_output_scale(self._handle)

    @output_scale.setter
    def output_scale(self, value):
        _check(self._lib.gcs_unet_set_output_scale(self._handle, float(value)),
               "gcs_unet_set_output_scale")

    def forward(self, features):
        features = _f32(features)
        batch = features.shape[0]
        out = np.zeros((batch, self.height, self.width), dtype=np.float32)
        _check(self._lib.gcs_unet_forward(self._handle, features.ctypes.data_as(_F32P),
                                          batch, out.ctypes.data_as(_F32P)),
               "gcs_unet_forward")
        return out

    def gradients(self, features, target):
        features, target = _f32(features), _f32(target)
        loss = ctypes.c_double()
        _check(self._lib.gcs_unet_gradients(self._handle, features.ctypes.data_as(_F32P),
                                            target.ctypes.data_as(_F32P),
                                            features.shape[0], ctypes.byref(loss)),
               "gcs_unet_gradients")
        return loss.value

    def adam_step(self, learning_rate):
        _check(self._lib.gcs_unet_adam_step(self._handle, float(learning_rate)),
               "gcs_unet_adam_step")


def gemm_check(a_bits, b_bits):
    """c = a @ b on the matrix cores; a, b given as uint16 bf16 patterns."""
    lib = library()
    _declare(lib)
    a_bits = np.ascontiguousarray(a_bits, dtype=np.uint16)
    b_bits = np.ascontiguousarray(b_bits, dtype=np.uint16)
    m, k = a_bits.shape
    n = b_bits.shape[1]
    c = np.zeros((m, n), dtype=np.float32)
    _check(lib.gcs_mfma_gemm_check(a_bits.ctypes.data, b_bits.ctypes.data, m, n, k,
                                   c.ctypes.data_as(_F32P)), "gcs_mfma_gemm_check")
    return c


def attach(problem, net):
    lib = problem._lib
    _declare(lib)
    _check(lib.gcs_problem_mi_attach_warm_start(problem._handle,
                                                net._handle if net is not None else None),
           "gcs_problem_mi_attach_warm_start")


def train(problem, net, steps, batch, learning_rate=1e-3, seed=1, pedersen=(5.0, 15.0),
          hall=(0.0, 10.0)):
    """Train `net` on generated pairs for `problem`'s MI grid; per-step losses."""
    lib = problem._lib
    _declare(lib)
    cfg = TrainConfig(steps, batch, learning_rate, seed, pedersen[0], pedersen[1], hall[0],
                      hall[1])
    losses = np.zeros(steps, dtype=np.float32)
    _check(lib.gcs_problem_mi_train_warm_start(problem._handle, net._handle,
                                               ctypes.byref(cfg),
                                               losses.ctypes.data_as(_F32P)),
           "gcs_problem_mi_train_warm_start")
    return losses


def train_on(problem, net, current, potential, hemisphere, steps, batch, learning_rate=1e-3,
             seed=1):
    """Train `net` on given pairs (current, potential [count][nt][np], each
    one hemisphere: GAMERA_MI_NORTH = -1 or GAMERA_MI_SOUTH = 1); per-step
    losses."""
    lib = problem._lib
    _declare(lib)
    current, potential = _f64(current), _f64(potential)
    hemi = np.ascontiguousarray(hemisphere, dtype=np.int32)
    count = current.shape[0]
    if potential.shape != current.shape or hemi.shape != (count,):
        raise ValueError("current, potential and hemisphere must describe the same pairs")
    cfg = TrainConfig(steps, batch, learning_rate, seed, 0.0, 0.0, 0.0, 0.0)
    losses = np.zeros(steps, dtype=np.float32)
    _check(lib.gcs_problem_mi_train_warm_start_on(
        problem._handle, net._handle, ctypes.byref(cfg), current.ctypes.data_as(_F64P),
        potential.ctypes.data_as(_F64P), hemi.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
        count, losses.ctypes.data_as(_F32P)), "gcs_problem_mi_train_warm_start_on")
    return losses


# Solar-wind ranges of the logged runs: Bz in [-10, 5] nT, n in [2, 20] cm^-3,
# V in [300, 700] km/s, uniform.
WIND_RANGES = {"bz_nt": (-10.0, 5.0), "n_cm3": (2.0, 20.0), "speed_km_s": (300.0, 700.0)}


def wind_variation(seed):
    """One seeded draw of the solar wind from WIND_RANGES."""
    rng = np.random.default_rng(seed)
    return {k: float(rng.uniform(lo, hi)) for k, (lo, hi) in WIND_RANGES.items()}


def log_currents(problem, samples, every, skip=0):
    """Advance `skip` driver steps, then `samples` times: `every` steps and a
    sample of the field-aligned current onto the MI grids. Returns
    [samples][2][nt][np] currents and their simulated times (code units)."""
    nt = problem.config.mi.colatitude_points
    np_ = problem.config.mi.azimuth_points
    if skip:
        problem.advance(skip)
    out = np.zeros((samples, 2, nt, np_))
    times = np.zeros(samples)
    for k in range(samples):
        problem.advance(every)
        problem.mi_sample_current()
        out[k] = problem.mi_field(1).reshape(2, nt, np_)
        times[k] = problem.clock()["time"]
    return out, times


def solve_targets(problem, currents):
    """The plain FP64 Krylov solve of each current [k][2][nt][np] to the
    problem's MI tolerance, each started from the previous solution; returns
    the potentials and the (north, south) iterations of each solve."""
    pots = np.zeros_like(currents)
    iters = []
    guess = np.zeros_like(currents[0])
    for k in range(currents.shape[0]):
        pots[k], it, _ = solve_current(problem, currents[k], guess, use_network=False)
        iters.append(it)
        guess = pots[k]
    return pots, iters


def hemisphere_pairs(currents, potentials, lag=0):
    """Training pairs from logged (current, potential) samples [k][2][nt][np]:
    lag 0 the samples themselves, lag L > 0 their differences L samples
    apart (what the warm start sees: the residual of the potential L
    samples earlier is the current's change, its error the potential's
    change). Each hemisphere is one pair."""
    if lag:
        currents = currents[lag:] - currents[:-lag]
        potentials = potentials[lag:] - potentials[:-lag]
    k, _, nt, np_ = currents.shape
    hemi = np.tile(np.array([-1, 1], dtype=np.int32), k)
    return currents.reshape(2 * k, nt, np_), potentials.reshape(2 * k, nt, np_), hemi


def make_pairs(problem, seed, count, pedersen, hall):
    """(current, potential), each [count][nt][np]; even samples north."""
    lib = problem._lib
    _declare(lib)
    nt = problem.config.mi.colatitude_points
    np_ = problem.config.mi.azimuth_points
    cur = np.zeros((count, nt, np_))
    pot = np.zeros((count, nt, np_))
    _check(lib.gcs_problem_mi_make_pairs(problem._handle, seed, count, pedersen, hall,
                                         cur.ctypes.data_as(_F64P),
                                         pot.ctypes.data_as(_F64P)),
           "gcs_problem_mi_make_pairs")
    return cur, pot


def solve_current(problem, current, guess, use_network):
    """(potential [2][nt][np], iterations (north, south), ms)."""
    lib = problem._lib
    _declare(lib)
    current = _f64(current)
    pot = _f64(guess).copy()
    iters = (ctypes.c_int * 2)()
    ms = ctypes.c_double()
    _check(lib.gcs_problem_mi_solve_current(problem._handle, current.ctypes.data_as(_F64P),
                                            pot.ctypes.data_as(_F64P),
                                            1 if use_network else 0, iters,
                                            ctypes.byref(ms)),
           "gcs_problem_mi_solve_current")
    return pot, (iters[0], iters[1]), ms.value


__all__ = ["WarmStartNet", "GpuProblem", "attach", "train", "train_on", "make_pairs",
           "solve_current", "log_currents", "solve_targets", "hemisphere_pairs",
           "wind_variation", "WIND_RANGES", "gemm_check", "IN_CHANNELS"]
