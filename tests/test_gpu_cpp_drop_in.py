"""Runs the C++ drop-in test programs (tests/cpp): lbm::SolverHip, SolverV3's method set over the
C ABI, against the oracle (tests/test_v3_vs_v0.cu, tests/test_cylinder_v3.cu shapes) and through
the reference's Poiseuille checks A/B/C (tests/poiseuille_check.hpp:81-158), in float and double
builds."""
from __future__ import annotations

import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = ROOT / "tests" / "cpp" / "bin"


@pytest.mark.parametrize("real", ("float", "double"))
def test_cpp_drop_in(real):
    exe = BIN / f"test_drop_in_{real}"
    assert exe.exists(), f"{exe} not built (build() compiles tests/cpp)"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


@pytest.mark.parametrize("real", ("float", "double"))
def test_cpp_poiseuille_drop_in(real):
    exe = BIN / f"test_poiseuille_{real}"
    assert exe.exists(), f"{exe} not built (build() compiles tests/cpp)"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
