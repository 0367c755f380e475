"""The hipGraph-captured driver step against the kernel-by-kernel step.

Two copies of one problem advance side by side, one launching every kernel
of the step, the other replaying the captured step graph. After every
replay the host downloads and compares the whole state (eager work and a
synchronisation between replays); the two must be equal bit for bit, also
across a precision switch and the ionosphere's first update, which both
force a recapture.
"""
import pytest

from conftest import bits_equal, gcs_config, problem

pytestmark = pytest.mark.gpu


def assert_same_state(a, b, where):
    for fid in problem.STATE_FIELDS:
        for f in range(gcs_config.FACE_COUNT):
            assert bits_equal(a.field(f, fid), b.field(f, fid)), (where, fid, f)


@pytest.mark.parametrize("name", ["small", "small_mi"])
def test_graph_replays_equal_eager_steps(gpu, name):
    cfg = gcs_config.workload(name)
    print(f"graph vs eager {name}", flush=True)
    a = gpu.GpuProblem(cfg)
    b = gpu.GpuProblem(cfg)
    b.set_graph(True)
    for s in range(8):
        if s == 5:  # recapture on a precision switch
            a.set_precision(True)
            b.set_precision(True)
        a.advance(1)
        b.advance(1)
        assert a.clock() == b.clock(), s
        assert_same_state(a, b, s)
    a.advance(3)
    b.advance(3)
    assert_same_state(a, b, "multi-step advance")
    if cfg.mi_enabled:
        assert a.mi_report()["updates"] == b.mi_report()["updates"] >= 1
