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
