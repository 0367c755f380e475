"""The oracle's full six-face step against the compiled reference library
(oracle/_ref), bitwise, on identical inputs."""
import numpy as np
import pytest

from conftest import (bits_equal, gcs_config, max_divergence, oracle_from,
                      problem)

CHECKED = (problem.STATE_FIELDS
           + [problem.PREDICTED_CONSERVED, problem.PREDICTED_CELL_MAGNETIC]
           + [problem.PREDICTED_FACE_FLUX + d for d in range(3)]
           + [problem.EDGE_EMF + d for d in range(3)]
           + [problem.FLUID_FACE_FLUX + d for d in range(3)]
           + [problem.MAXWELL_FACE_FLUX + d for d in range(3)])


@pytest.mark.parametrize("name,steps", [("tiny", 4), ("small", 12), ("ragged", 6)])
def test_oracle_steps_bitwise_equal_reference(ref, orc, name, steps):
    cfg = gcs_config.workload(name)
    r = problem.Problem(ref, "refh", cfg)
    o = oracle_from(r, orc, cfg)
    for _ in range(steps):
        r.advance(1)
        o.advance(1)
        assert r.clock() == o.clock()
        for fid in CHECKED:
            for f in range(gcs_config.FACE_COUNT):
                assert bits_equal(r.field(f, fid), o.field(f, fid)), (fid, f)
    # the reference's own gate: positive state, div B at roundoff
    assert max_divergence(o, cfg) < 1e-9
    for f in range(gcs_config.FACE_COUNT):
        u = o.field(f, problem.CONSERVED).reshape(-1, 5)
        assert np.all(u[:, 0] > 0)


def test_oracle_explicit_step_matches_reference(ref, orc):
    cfg = gcs_config.workload("small")
    r = problem.Problem(ref, "refh", cfg)
    o = oracle_from(r, orc, cfg)
    for ratio, dt in ((0.0, 2e-3), (0.37, 3e-3)):
        r.step_explicit(ratio, dt, 0.1, 0.097)
        o.step_explicit(ratio, dt, 0.1, 0.097)
        for fid in problem.STATE_FIELDS:
            for f in range(gcs_config.FACE_COUNT):
                assert bits_equal(r.field(f, fid), o.field(f, fid)), (fid, f)
