"""The 1024x256 force-driven cylinder instance (D 32 at (256, 129), Re 100, 40,000 steps) on the
GPU: the float run's wake-probe signal is bit-identical to the oracle's, and the Strouhal
numbers of the float, double and FP16-storage runs agree.

The probe is u_y at (xc + 3D, yc) every 20 steps, as the reference's cylinder driver samples it
(src/common/driver.hpp:175-214); the Strouhal analysis follows scripts/strouhal.py:57-81 (FFT
peak and upward zero crossings over the second half of the run, which must agree within 0.01).
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def runs(lbm):
    return {prec: lbm.cylinder_instance(prec) for prec in ("float", "double", "fp16")}


def test_float_probe_signal_bitwise_equals_oracle(lbm, runs):
    g = np.load(GOLDEN / "cylinder_probe_1024x256_40000_float.npz")
    p = json.loads(str(g["params"]))
    r = runs["float"]
    assert abs(r["tau"] - p["tau"]) == 0 and abs(r["force"] - p["force"]) == 0
    probe = g["probe"]
    assert r["probe"].shape == probe.shape
    mism = np.flatnonzero(r["probe"] != probe)
    assert mism.size == 0, f"first differing sample {mism[0]} of {probe.size}"


def test_strouhal_methods_agree_and_precisions_match(lbm, runs):
    for prec, r in runs.items():
        print(f"{prec}: St {r['st_fft']:.4f} (FFT) / {r['st_zero']:.4f} (zero crossings) over "
              f"{r['periods']} periods, lift St {r['st_lift_fft']:.4f}, Re_w {r['Re_window']:.1f}")
        assert r["periods"] >= 5
        assert abs(r["st_fft"] - r["st_zero"]) <= 0.01  # scripts/strouhal.py:81
    st64 = runs["double"]["st_fft"]
    assert abs(runs["float"]["st_fft"] - st64) / st64 < 1e-3
    err16 = abs(runs["fp16"]["st_fft"] - st64) / st64
    print(f"fp16 Strouhal relative error vs fp64: {err16:.2e}")
    assert err16 < 1e-2
