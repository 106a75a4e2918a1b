"""The 1024x256 force-driven cylinder instance (D 32 at (256, 129), Re 100, 40,000 steps) on the
GPU: the body force found by cylinder_force() holds the window Reynolds number at 100 (1 %), the
float run's wake-probe signal is bit-identical to the oracle's at that force, the Strouhal number
from the lift agrees with the probe's, and the float, double and FP16-storage runs agree.

The probe is u_y at (xc + 3D, yc) every 20 steps, as the reference's cylinder driver samples it
(src/common/driver.hpp:175-214); the Strouhal analysis follows scripts/strouhal.py:57-81 (FFT
peak and upward zero crossings over the second half of the run, which must agree within 0.01);
the velocity scale is the window mean of the fluid velocity (FP64 momentum reduction).
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def calibrated(lbm):
    print("cylinder force calibration (FP64)", flush=True)
    return lbm.cylinder_force()


@pytest.fixture(scope="module")
def runs(lbm, calibrated):
    force = calibrated[0]
    out = {}
    for prec in ("float", "double", "fp16"):
        print(f"cylinder instance {prec}", flush=True)
        out[prec] = lbm.cylinder_instance(prec, force=force)
    return out


def test_calibrated_force_holds_re_100(lbm, calibrated):
    force, uw, cal_runs = calibrated
    print(f"force {force!r}, window velocity {uw:.6f}, runs {cal_runs}")
    assert abs(uw / 0.05 - 1.0) < 2e-3
    g = json.loads(str(np.load(GOLDEN / "cylinder_probe_1024x256_40000_float.npz")["params"]))
    # the fixture was written at the force this calibration returns (deterministic FP64 runs)
    assert abs(force / g["force"] - 1.0) < 1e-6


def test_float_probe_signal_bitwise_equals_oracle(lbm, runs):
    g = np.load(GOLDEN / "cylinder_probe_1024x256_40000_float.npz")
    p = json.loads(str(g["params"]))
    r = lbm.cylinder_instance("float", force=p["force"])
    assert r["tau"] == p["tau"]
    probe = g["probe"]
    assert r["probe"].shape == probe.shape
    mism = np.flatnonzero(r["probe"] != probe)
    assert mism.size == 0, f"first differing sample {mism[0]} of {probe.size}"


def test_reynolds_and_strouhal(lbm, runs):
    for prec, r in runs.items():
        print(f"{prec}: Re_w {r['Re_window']:.2f}; lift St {r['st_lift_fft']:.4f} (FFT) / "
              f"{r['st_lift_zero']:.4f} (zero crossings, {r['lift_periods']} periods); probe St "
              f"{r['st_fft']:.4f} / {r['st_zero']:.4f} ({r['periods']} periods)")
        assert abs(r["Re_window"] - 100.0) <= 1.0
        assert r["periods"] >= 5 and r["lift_periods"] >= 5
        assert abs(r["st_fft"] - r["st_zero"]) <= 0.01  # scripts/strouhal.py:81
        assert abs(r["st_lift_fft"] - r["st_lift_zero"]) <= 0.01
        # the lift and the wake probe shed at the same frequency
        assert abs(r["st_lift_fft"] - r["st_fft"]) / r["st_fft"] < 0.01
    st64 = runs["double"]["st_lift_fft"]
    assert abs(runs["float"]["st_lift_fft"] - st64) / st64 < 2e-3
    err16 = abs(runs["fp16"]["st_lift_fft"] - st64) / st64
    print(f"fp16 lift Strouhal relative error vs fp64: {err16:.2e}")
    assert err16 < 1e-2
