"""The serial Earth driver's CSV output
(cubedsphere/tests/test_cubedsphere_dipole_background.c:120-193): the
six-face snapshot and the output frame with its index line.

The reference's own writers are compiled from the driver's source into the
harness library (oracle/ref_snapshot.c). The oracle's restatement
(orc_write_snapshot) must write the same bytes on the same state (CPU), and
libgcs.so's device-computed snapshot must write the same bytes as the
reference's writer on the GPU problem's state copied into a reference
problem (GPU)."""
import ctypes
import os

import pytest

from conftest import REF_LIB, gcs_config, oracle_from, problem, reference_from

CP = ctypes.c_char_p
VP = ctypes.c_void_p


def declare_ref(ref):
    ref.refh_write_snapshot.argtypes = [VP, CP]
    ref.refh_write_drive_frame.argtypes = [VP, CP, ctypes.c_long, ctypes.c_long, ctypes.c_double,
                                           CP]


def read(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.fixture(scope="module")
def ref_lib():
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    lib = ctypes.CDLL(REF_LIB)
    declare_ref(lib)
    return lib


def test_oracle_snapshot_is_the_drivers_bytes(ref_lib, orc, tmp_path):
    cfg = gcs_config.workload("small")
    r = problem.Problem(ref_lib, "refh", cfg)
    r.advance(3)
    o = oracle_from(r, orc, cfg)
    orc.orc_write_snapshot.argtypes = [VP, CP]
    a, b = tmp_path / "ref.csv", tmp_path / "orc.csv"
    assert ref_lib.refh_write_snapshot(r._handle, str(a).encode()) == 0
    assert orc.orc_write_snapshot(o._handle, str(b).encode()) == 0
    text = read(a)
    lines = text.splitlines()
    assert lines[0] == b"face,i,j,k,x,y,z,rho,p,vx,vy,vz,bx,by,bz"
    assert len(lines) == 1 + cfg.active_cells
    assert lines[1].startswith(b"0,0,0,0,") and lines[-1].startswith(
        f"5,{cfg.radial_cells - 1},{cfg.angular_cells - 1},{cfg.angular_cells - 1},".encode())
    assert read(b) == text


def test_drive_frames_and_index(ref_lib, tmp_path):
    """Two frames and their index lines (:177-193, header from :561)."""
    cfg = gcs_config.workload("tiny")
    r = problem.Problem(ref_lib, "refh", cfg)
    prefix, index = tmp_path / "drive", tmp_path / "drive.index.csv"
    for frame, step, seconds in ((0, 1, 0.5), (1, 7, 600.25)):
        r.advance(1)
        assert ref_lib.refh_write_drive_frame(r._handle, str(prefix).encode(), frame, step,
                                              seconds, str(index).encode()) == 0
    lines = read(index).decode().splitlines()
    assert lines == ["frame,step,simulation_seconds,mhd_csv",
                     f"0,1,0.5,{prefix}_000000.csv", f"1,7,600.25,{prefix}_000001.csv"]


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps", [("small", 6), ("ragged", 4)])
def test_device_snapshot_is_the_drivers_bytes(gpu, name, steps, tmp_path):
    """gcs_problem_write_snapshot (primitives and total B computed on the
    device, formatted on the host) against the driver's writer on the same
    state, byte for byte; and the drive frame's CSV and index line."""
    if not os.path.exists(REF_LIB):
        pytest.skip("compiled reference (oracle/_ref) not built")
    ref = ctypes.CDLL(REF_LIB)
    declare_ref(ref)
    cfg = gcs_config.workload(name)
    g = gpu.GpuProblem(cfg)
    g.advance(steps)
    lib = gpu.library()
    lib.gcs_problem_write_snapshot.argtypes = [VP, CP]
    lib.gcs_problem_write_drive_frame.argtypes = [VP, CP, ctypes.c_long, ctypes.c_long,
                                                  ctypes.c_double, CP]
    ours, theirs = tmp_path / "gcs.csv", tmp_path / "ref.csv"
    assert lib.gcs_problem_write_snapshot(g._handle, str(ours).encode()) == 0, gpu.last_error()
    r = reference_from(g, ref, cfg)
    assert ref.refh_write_snapshot(r._handle, str(theirs).encode()) == 0
    a, b = read(ours), read(theirs)
    lines = a.count(b"\n")
    print(f"snapshot {name}: {len(a)} bytes, {lines} lines", flush=True)
    assert a == b
    # an output frame: the same CSV under the frame name and its index line
    prefix, index = tmp_path / "frame", tmp_path / "frame.index.csv"
    assert lib.gcs_problem_write_drive_frame(g._handle, str(prefix).encode(), 3, steps, 1.25,
                                             str(index).encode()) == 0
    assert read(f"{prefix}_000003.csv") == b
    assert read(index).decode().splitlines() == [
        "frame,step,simulation_seconds,mhd_csv", f"3,{steps},1.25,{prefix}_000003.csv"]
