"""CPU check of the device field arithmetic (truepix_amd/csrc/fp61.hpp).

The header's integer code is what the GPU runs; built with plain g++ it is
checked against 128-bit products: the split-31 multiply at the operands that
set its no-overflow bounds, the fold, and the lazy Fp2 dot product against an
exact one (tests/field_check.cpp).  The field follows
TP/native_fp2/fp61.cuh:14-166 (p = 2^61 - 1, i^2 = -1).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_field_arithmetic(tmp_path):
    exe = str(tmp_path / "field_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "truepix_amd", "csrc"),
                    os.path.join(ROOT, "tests", "field_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_claim_terms_equal_the_gate_sum(tmp_path):
    """The early claim's term table (csrc/claim_terms.hpp) gives the same four
    claim values as the gate-by-gate sum, for layers over all 13 gate types."""
    exe = str(tmp_path / "claim_terms_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "truepix_amd", "csrc"),
                    os.path.join(ROOT, "tests", "claim_terms_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
