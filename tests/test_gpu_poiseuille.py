"""Poiseuille known-answer checks on the GPU step (reference tests/poiseuille_check.hpp:33-158).

Channel periodic in x (nx = 4), halfway bounce-back walls (H = ny - 2), body force from
poiseuille_force(H, tau, u_max = 0.05). Checks:
  A. tau = 1/2 + sqrt(3)/4 (BGK + halfway bounce-back exact): L2 < 1e-8 (double), 2e-4 (float);
  B. tau = 1, H = 16, 32, 64: L2 < 2e-3; in double the observed order is in (1.9, 2.1);
  C. relative mass drift < 1e-13 (double), 1e-6 (float).
Steady state: relative change of the x-averaged profile over 1000 steps below 1e-10 (double) /
1e-7 (float), capped at 20 H^2 / nu steps."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"double": dict(steady=1e-10, exact=1e-8, err=2e-3, mass=1e-13),
       "float": dict(steady=1e-7, exact=2e-4, err=2e-3, mass=1e-6)}
UMAX, NX, EVERY = 0.05, 4, 1000


def profile(s, ny):
    ux = s.macroscopic()[1].reshape(ny, NX).astype(np.float64)
    return ux[1:ny - 1].sum(axis=1) / NX


def rel_l2(a, b):
    return math.sqrt(float(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def run_poiseuille(lbm, H, tau, prec):
    ny = H + 2
    tol = TOL[prec]
    force = lbm.poiseuille_force(H, tau, UMAX, prec)
    nu = float(np.float32((np.float32(tau) - np.float32(0.5)) / np.float32(3))) \
        if prec == "float" else (tau - 0.5) / 3.0
    s = lbm.Solver(NX, ny, tau, force, 0.0, lbm.make_channel(NX, ny), prec)
    s.init_equilibrium(1.0, 0.0, 0.0)
    mass0 = s.total_mass()
    max_steps = int(20.0 * H * H / nu)
    prev = profile(s, ny)
    steps, steady = 0, False
    while steps < max_steps:
        s.run(EVERY)
        steps += EVERY
        cur = profile(s, ny)
        done = rel_l2(cur, prev) < tol["steady"]
        prev = cur
        if done:
            steady = True
            break
    k = np.arange(H) + 0.5
    exact = force / (2.0 * nu) * k * (H - k)
    return rel_l2(prev, exact), (s.total_mass() - mass0) / mass0, steps, steady


@pytest.mark.parametrize("prec", ("float", "double"))
def test_exact_tau(lbm, prec):
    tau = 0.5 + math.sqrt(3.0) / 4.0
    l2, drift, steps, steady = run_poiseuille(lbm, 32, tau, prec)
    print(f"A {prec}: L2={l2:.3e} steps={steps} steady={steady} drift={drift:.2e}")
    assert l2 < TOL[prec]["exact"]
    assert abs(drift) < TOL[prec]["mass"]


@pytest.mark.parametrize("prec", ("float", "double"))
def test_tau_one_convergence(lbm, prec):
    errs = []
    for H in (16, 32, 64):
        l2, drift, steps, steady = run_poiseuille(lbm, H, 1.0, prec)
        print(f"B {prec} H={H}: L2={l2:.3e} steps={steps} steady={steady} drift={drift:.2e}")
        assert l2 < TOL[prec]["err"]
        assert abs(drift) < TOL[prec]["mass"]
        errs.append(l2)
    if prec == "double":
        for a, b in zip(errs, errs[1:]):
            order = math.log2(a / b)
            assert 1.9 < order < 2.1, order
