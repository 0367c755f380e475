"""The shared-reciprocal division of the face flux (div_by_reciprocal in
gamera-cube_amd/csrc/gcs_math.hpp) is bit for bit IEEE division: the same
three-operation sequence in C with fused multiply-adds, and the plain
division outside its exact range, on 2e7 random normal operands with crowded
significands and on zeros, infinities, NaN, subnormal and near-overflow
operands (tests/c/reciprocal_division.c)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reciprocal_correction_is_ieee_division(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = tmp_path / "reciprocal_division"
    # glibc's fma() is correctly rounded with or without the x86 FMA unit,
    # like the GPU's fused instruction; no target flag is needed
    subprocess.run([cc, "-O2", "-ffp-contract=off", "-fno-builtin", "-o", str(exe),
                    os.path.join(HERE, "c", "reciprocal_division.c"), "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    print(out)
    assert "mismatches=0" in out
