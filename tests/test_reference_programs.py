"""The reference's own step tests and Earth driver, linked against
libgcs.so in place of cubedsphere_step.c (oracle/Makefile `programs`).

Each program keeps its own setup, callbacks and pass/fail gates; only the
six-face step runs on the GPU. It must exit as the reference build does,
and every number it prints must equal the reference build's printed value
(tests/golden/programs.json, written by make_program_golden.py) to the
printed precision. Quantities at roundoff level (div B, recovery error,
symmetry defects below 1e-9) only have to stay at roundoff.
"""
import json
import os
import re
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "oracle", "_ref", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden", "programs.json")
NUMBER = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")

with open(GOLDEN) as f:
    CASES = json.load(f)


def numbers(text):
    return [float(x) for x in NUMBER.findall(text)]


def same_printed(a, b):
    if abs(a) < 1e-9 and abs(b) < 1e-9:
        return True  # roundoff-level diagnostics
    # %e with 7 significant digits (and %g with 6): one unit in the last
    # printed place, relative.
    return abs(a - b) <= 2e-6 * max(abs(a), abs(b))


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_program_with_dropin_step(case):
    spec = CASES[case]
    exe = os.path.join(BIN, f"test_cubedsphere_{spec['program']}_gcs")
    if not os.path.exists(exe):
        pytest.skip("reference programs not built (make -C oracle programs)")
    print(f"program {case}: {spec['program']} {' '.join(spec['args'])}", flush=True)
    run = subprocess.run([exe, *spec["args"]], capture_output=True, text=True,
                         timeout=240)
    print(run.stdout[-2000:], run.stderr[-2000:], flush=True)
    assert run.returncode == spec["rc"], run.stderr[-1000:]
    got, want = numbers(run.stdout), numbers(spec["stdout"])
    assert len(got) == len(want), (run.stdout, spec["stdout"])
    bad = [(g, w) for g, w in zip(got, want) if not same_printed(g, w)]
    assert not bad, bad
