"""ctypes wrapper over the `<prefix>_create/_advance/_field` problem API.

The HIP library exports this API with prefix `gcs_problem`; the CPU oracle
and the compiled-reference harness export the same shape under their own
prefixes, which lets tests drive all three through one class.
"""
import ctypes

import numpy as np

from .config import FACE_COUNT, ProblemConfig

# Field ids (include/gamera_cs_problem.h).
CONSERVED = 0
OLD_CONSERVED = 1
PREDICTED_CONSERVED = 2
CELL_MAGNETIC = 3
OLD_CELL_MAGNETIC = 4
PREDICTED_CELL_MAGNETIC = 5
FACE_FLUX = 6
OLD_FACE_FLUX = 9
PREDICTED_FACE_FLUX = 12
EDGE_EMF = 15
FLUID_FACE_FLUX = 18
MAXWELL_FACE_FLUX = 21
HYDRO_RATE = 24
MAXWELL_RATE = 25
VERTEX = 26
CELL_GEOMETRY = 27
FACE_GEOMETRY = 28
EDGE_GEOMETRY = 31
BG_CELL_MAGNETIC = 34
BG_CELL_FORCE = 35
BG_FACE_MAGNETIC = 36
BG_FACE_FLUX = 39
BG_EDGE_MAGNETIC = 42
FIELD_COUNT = 45

STATE_FIELDS = ([CONSERVED, OLD_CONSERVED, CELL_MAGNETIC, OLD_CELL_MAGNETIC]
                + [FACE_FLUX + d for d in range(3)]
                + [OLD_FACE_FLUX + d for d in range(3)])
GEOMETRY_FIELDS = ([VERTEX, CELL_GEOMETRY]
                   + [FACE_GEOMETRY + d for d in range(3)]
                   + [EDGE_GEOMETRY + d for d in range(3)])
BACKGROUND_FIELDS = ([BG_CELL_MAGNETIC, BG_CELL_FORCE]
                     + [BG_FACE_MAGNETIC + d for d in range(3)]
                     + [BG_FACE_FLUX + d for d in range(3)]
                     + [BG_EDGE_MAGNETIC + d for d in range(3)])


def bind(lib, prefix, create="create"):
    """Declare argument/return types of one problem API on a CDLL."""
    def f(name, restype, *argtypes):
        fn = getattr(lib, f"{prefix}_{name}")
        fn.restype = restype
        fn.argtypes = list(argtypes)
        return fn
    vp = ctypes.c_void_p
    return {
        "create": f(create, vp, ctypes.POINTER(ProblemConfig)),
        "destroy": f("destroy", None, vp),
        "advance": f("advance", ctypes.c_int, vp, ctypes.c_long),
        "step_explicit": f("step_explicit", ctypes.c_int, vp, ctypes.c_double,
                           ctypes.c_double, ctypes.c_double, ctypes.c_double),
        "clock": f("clock", None, vp, ctypes.POINTER(ctypes.c_double)),
        "field": f("field", ctypes.c_int, vp, ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                   ctypes.POINTER(ctypes.c_size_t)),
        "edge_valid": f("edge_valid", ctypes.c_int, vp, ctypes.c_int,
                        ctypes.c_int,
                        ctypes.POINTER(ctypes.POINTER(ctypes.c_ubyte)),
                        ctypes.POINTER(ctypes.c_size_t)),
    }


class Problem:
    """One six-face problem owned by a native library."""

    def __init__(self, lib, prefix, config, create="create"):
        """`create` names the constructor: `<prefix>_create` sets the
        problem up; the CPU implementations also export constructors that
        only allocate (the caller copies the arrays in)."""
        if not isinstance(config, ProblemConfig):
            raise TypeError("config must be a ProblemConfig")
        self._api = bind(lib, prefix, create)
        self._lib = lib
        self.prefix = prefix
        self.config = config
        self._handle = self._api["create"](ctypes.byref(config))
        if not self._handle:
            raise RuntimeError(f"{prefix}_create failed")

    def close(self):
        if self._handle:
            self._api["destroy"](self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, steps):
        status = self._api["advance"](self._handle, int(steps))
        if status != 0:
            raise RuntimeError(f"{self.prefix}_advance failed ({status})")

    def step_explicit(self, ratio, dt, time, previous_time):
        status = self._api["step_explicit"](self._handle, ratio, dt, time,
                                            previous_time)
        if status != 0:
            raise RuntimeError(f"{self.prefix}_step_explicit failed")

    def clock(self):
        out = (ctypes.c_double * 4)()
        self._api["clock"](self._handle, out)
        return {"time": out[0], "dt": out[1], "previous_dt": out[2],
                "steps": int(out[3])}

    def field(self, face, field_id):
        """Zero-copy numpy view of one face's array (host memory)."""
        ptr = ctypes.POINTER(ctypes.c_double)()
        count = ctypes.c_size_t()
        if self._api["field"](self._handle, face, field_id, ctypes.byref(ptr),
                              ctypes.byref(count)) != 0:
            raise KeyError(field_id)
        return np.ctypeslib.as_array(ptr, shape=(count.value,))

    def edge_valid(self, face, direction):
        ptr = ctypes.POINTER(ctypes.c_ubyte)()
        count = ctypes.c_size_t()
        if self._api["edge_valid"](self._handle, face, direction,
                                   ctypes.byref(ptr),
                                   ctypes.byref(count)) != 0:
            raise KeyError(direction)
        return np.ctypeslib.as_array(ptr, shape=(count.value,))

    def copy_fields(self, other, field_ids):
        """Overwrite this problem's arrays with another problem's."""
        for face in range(FACE_COUNT):
            for fid in field_ids:
                dst = self.field(face, fid)
                src = other.field(face, fid)
                if dst.shape != src.shape:
                    raise ValueError(f"field {fid} shape mismatch")
                dst[:] = src
