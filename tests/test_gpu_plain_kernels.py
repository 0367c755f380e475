"""The LDS-tiled face-flux and edge-EMF kernels and the register-march
face-velocity kernels against the one-face-per-thread kernels, and the
seam recovery's gathered data and the cached ionosphere drift against
their reads from the planes and the per-step drift: the same expressions
per value, so every array must be bit-identical, in FP64 and in the mixed
form, at the ragged parity size (every tile partial), at the small size
(16 x 12^2 per face), with the MI coupling across two ionosphere updates,
and at config 1 (6 x 96^2 x 192), the size the tiles were tuned at."""
import pytest

from conftest import bits_equal, gcs_config, problem

pytestmark = pytest.mark.gpu

CHECKED = (problem.STATE_FIELDS
           + [problem.PREDICTED_CONSERVED, problem.PREDICTED_CELL_MAGNETIC]
           + [problem.EDGE_EMF + d for d in range(3)]
           + [problem.FLUID_FACE_FLUX + d for d in range(3)]
           + [problem.MAXWELL_FACE_FLUX + d for d in range(3)])


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,steps,mixed", [("ragged", 6, False), ("ragged", 6, True),
                                              ("small", 4, False),
                                              ("small_mi", 44, False),
                                              ("config1_nomi", 2, False),
                                              ("config1_nomi", 2, True)])
def test_tiled_kernels_bitwise_equal_plain_kernels(gpu, name, steps, mixed):
    cfg = gcs_config.workload(name)
    print(f"tiled vs plain kernels: {name} x{steps} {'mixed' if mixed else 'FP64'}",
          flush=True)
    a = gpu.GpuProblem(cfg)
    b = gpu.GpuProblem(cfg)
    b.set_plain_kernels(True)
    if mixed:
        a.set_precision(True)
        b.set_precision(True)
    for s in range(steps):
        a.advance(1)
        b.advance(1)
        assert a.clock() == b.clock()
    for fid in CHECKED:
        for f in range(gcs_config.FACE_COUNT):
            assert bits_equal(a.field(f, fid), b.field(f, fid)), (fid, f)
