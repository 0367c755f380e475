"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: nothing here calls into the device)."""
import ctypes
import glob
import os
import re

from conftest import ROOT

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(g[a-z_0-9]+)\s*\([^;{]*\)\s*;", text)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    for n in ("gcs_problem_create", "gamera_cs_advance_six_with_background_planned",
              "gamera_cs_mpi_advance_six_with_background_planned", "gcs_unet_forward"):
        assert n in names


def test_library_exports_every_declared_symbol(gpu):
    lib = gpu.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_config_struct_matches_library(gpu):
    lib = gpu.library()
    assert lib.gcs_config_size() == ctypes.sizeof(gpu.ProblemConfig)


def test_invalid_config_is_rejected_before_any_device_work(gpu):
    cfg = gpu.ProblemConfig()  # zero ghost / zero cells
    lib = gpu.library()
    lib.gcs_problem_create.restype = ctypes.c_void_p
    lib.gcs_problem_create.argtypes = [ctypes.POINTER(gpu.ProblemConfig)]
    assert not lib.gcs_problem_create(ctypes.byref(cfg))
    assert "ghost" in gpu.last_error()
