"""The CPU baseline's protocol runner (oracle/cpu_protocol.py, run by bench.py's cpu_baseline leg):
its seeded-block instance is the package's (the GPU protocol's 2048^2 instance), and one run
reports the reference protocol's statistics (src/common/driver.hpp:115-132,
bench/run_bench.py:138-139: warm-up steps / 5, 5 timed repetitions, mean / std / max)."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

from conftest import ROOT


def test_seeded_block_is_the_package_instance(lbm):
    from oracle import cpu_protocol

    for seed, n in ((0, 2048), (1, 53), (5, 300)):
        assert (cpu_protocol.seeded_block(seed, n, n) == lbm.seeded_block_mask(seed, n, n)).all()


def test_protocol_runner_reports_five_repetitions(oracle_mod):
    env = dict(os.environ, OMP_NUM_THREADS="2", OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_protocol", "--nx", "96", "--ny", "64",
                        "--steps", "25", "--mask", "block", "--tau", "0.8", "--fx", "1e-5",
                        "--fy", "2e-6"], cwd=str(ROOT), env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    m = np.array(res["mlups"])
    assert res["reps"] == 5 and m.size == 5 and res["warmup"] == 5 and res["steps"] == 25
    assert res["threads"] == 2
    assert res["kind"] == ("reference" if oracle_mod.Reference.available("float") else "port")
    assert np.isclose(res["mean"], m.mean(), rtol=1e-3) and np.isclose(res["max"], m.max())
    assert np.isclose(res["std"], m.std(), rtol=1e-2, atol=0.02)
    secs = np.array(res["seconds"])
    assert np.allclose(96 * 64 * 25 / secs / 1e6, m, rtol=1e-2)
