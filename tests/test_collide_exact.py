"""The library's collision (csrc/d2q9.hpp::collide) against the reference's collision written
out in full, bit for bit, in float and double, on the CPU (tests/cpp/test_collide.cpp; built by
__graft_entry__.build()). The library leaves out products and sums that cannot change a finite
result; this pins that claim on random, signed-zero and subnormal inputs."""
from __future__ import annotations

import subprocess

from conftest import ROOT

EXE = ROOT / "tests" / "cpp" / "bin" / "test_collide"


def test_collide_equals_the_reference_expressions():
    assert EXE.exists(), f"{EXE} not built (build() compiles tests/cpp)"
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    for line in r.stdout.splitlines():
        if "cases compared" in line:
            assert int(line.split(":")[1].split()[0]) > 1_000_000, line
