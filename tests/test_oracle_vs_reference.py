"""The oracle's full six-face step against the compiled reference library
(oracle/_ref), bitwise, on identical inputs."""
import numpy as np
import pytest

from conftest import (bits_equal, counters, gcs_config, max_divergence,
                      oracle_from, problem, reference_from, set_wind, wind_series)

CHECKED = (problem.STATE_FIELDS
           + [problem.PREDICTED_CONSERVED, problem.PREDICTED_CELL_MAGNETIC]
           + [problem.PREDICTED_FACE_FLUX + d for d in range(3)]
           + [problem.EDGE_EMF + d for d in range(3)]
           + [problem.FLUID_FACE_FLUX + d for d in range(3)]
           + [problem.MAXWELL_FACE_FLUX + d for d in range(3)])


@pytest.mark.parametrize("name,steps", [("tiny", 4), ("small", 12), ("ragged", 6),
                                        ("fast_wind", 5), ("ghost5", 4)])
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
        # emergency-diffusion counters (nonorthogonal_sweep.c:541-545)
        for f in range(gcs_config.FACE_COUNT):
            assert bits_equal(counters(r, f), counters(o, f)), f
    fired = sum(counters(o, f)[0] for f in range(gcs_config.FACE_COUNT))
    if name == "fast_wind":
        print(f"fast_wind: emergency diffusion on {fired:.0f} faces in {steps} steps")
        assert fired > 0
    else:
        assert fired == 0
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


def test_reference_shell_steps_like_reference(ref):
    """refh_create_shell + copied arrays (how the CPU baseline times the
    reference at sizes whose reference setup takes minutes) steps bit for
    bit like the reference problem it was copied from, from mid-run."""
    cfg = gcs_config.workload("small")
    r = problem.Problem(ref, "refh", cfg)
    r.advance(2)
    s = reference_from(r, ref, cfg)
    for _ in range(3):
        r.advance(1)
        s.advance(1)
        assert r.clock() == s.clock()
        for fid in problem.STATE_FIELDS:
            for f in range(gcs_config.FACE_COUNT):
                assert bits_equal(r.field(f, fid), s.field(f, fid)), (fid, f)


def test_oracle_config0_200_steps_bitwise_equal_reference(ref, orc):
    """Config 0 (6 x 24^2 x 48) for its 200 steps: the oracle is the
    compiled reference bit for bit over the whole run, through the halo
    positivity fallback the run first takes at step 141."""
    cfg = gcs_config.workload("config0")
    r = problem.Problem(ref, "refh", cfg)
    o = oracle_from(r, orc, cfg)
    for chunk in range(4):
        for _ in range(50):
            r.advance(1)
            o.advance(1)
            assert r.clock() == o.clock()
        for fid in CHECKED:
            for f in range(gcs_config.FACE_COUNT):
                assert bits_equal(r.field(f, fid), o.field(f, fid)), (chunk, fid, f)
    assert max_divergence(o, cfg) < 1e-9


@pytest.mark.parametrize("linear", [True, False])
def test_oracle_long_wind_series_bitwise_equal_reference(ref, orc, linear):
    """A 40-sample wind series (the reference accepts any length,
    solar_wind.c:55-80): the Earth boundary's ballistic sampling, wind
    weights and outer EMF over a run, oracle and reference bit for bit."""
    cfg = gcs_config.workload("small")
    r = problem.Problem(ref, "refh", cfg)
    o = oracle_from(r, orc, cfg)
    time, state = wind_series()
    set_wind(r, time, state, linear)
    set_wind(o, time, state, linear)
    for _ in range(6):
        r.advance(1)
        o.advance(1)
        assert r.clock() == o.clock()
        for fid in problem.STATE_FIELDS:
            for f in range(gcs_config.FACE_COUNT):
                assert bits_equal(r.field(f, fid), o.field(f, fid)), (fid, f)
