"""Pins the oracle (CPU restatement of the reference solver) before it is trusted as a checker.

  * golden fixtures: outputs of the reference's own v0, compiled from its sources
    (tests/golden/make_golden.py), must be reproduced bit for bit, float and double;
  * closed-form identities of the equilibrium and the Guo source (moments), collision fixed
    points, momentum gain F per fluid cell per step in a periodic box;
  * Poiseuille exactness at tau = 1/2 + sqrt(3)/4 and second-order convergence at tau = 1
    (double), mass conservation;
  * the OpenMP pull form equals the sequential push form for 1, 2, 3, 4, 8 threads
    (reference tests/test_v1_vs_v0.cpp:47-64);
  * pull preimage / gather are inverse on fluid slots.
A plausible mistake (dropped term, wrong sign, wrong direction index, transposed bounce-back)
fails at least one of these."""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, walls_block_mask


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("kind", ("v0", "pull"))
def test_walls_block_golden(oracle_mod, prec, kind):
    solid, ref, p = load_golden(f"walls_block_53x31_500_{prec}.npz")
    o = oracle_mod.Oracle(53, 31, p["tau"], p["fx"], p["fy"], solid, prec, kind)
    o.init_equilibrium(*p["init"])
    o.run(p["steps"])
    np.testing.assert_array_equal(bits(o.populations()), bits(ref))


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("kind", ("v0", "pull"))
def test_cylinder_golden(oracle_mod, prec, kind):
    solid, ref, p = load_golden(f"cylinder_120x48_1000_{prec}.npz")
    np.testing.assert_array_equal(solid, oracle_mod.make_cylinder(p["nx"], p["ny"], p["xc"], p["yc"],
                                                                  p["D"]))
    o = oracle_mod.Oracle(p["nx"], p["ny"], p["tau"], 0.0, 0.0, solid, prec, kind)
    o.set_inflow_outflow(p["u_in"])
    o.init_equilibrium(*p["init"])
    o.run(p["steps"])
    np.testing.assert_array_equal(bits(o.populations()), bits(ref))


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("seed", (0, 1, 2))
def test_seeded_block_golden(oracle_mod, prec, seed):
    solid, ref, p = load_golden(f"seeded_block_53x31_500_seed{seed}_{prec}.npz")
    x0, y0 = p["block"]
    m = solid.reshape(31, 53)
    assert m[y0:y0 + 8, x0:x0 + 8].all() and m.sum() == 2 * 53 + 64
    o = oracle_mod.Oracle(53, 31, p["tau"], p["fx"], p["fy"], solid, prec, "v0")
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.run(p["steps"])
    np.testing.assert_array_equal(bits(o.populations()), bits(ref))


def test_crop256_double_golden_hash(oracle_mod):
    import hashlib

    h = json.loads((GOLDEN / "hashes.json").read_text())["crop256_f64"]
    o = oracle_mod.Oracle(256, 256, h["tau"], h["fx"], h["fy"], oracle_mod.make_channel(256, 256),
                          "double", "v0")
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.run(h["steps"])
    S = o.populations()
    assert hashlib.sha256(S.tobytes()).hexdigest() == h["sha256"]
    assert o.total_mass() == h["total_mass"]


@pytest.mark.parametrize("threads", (1, 2, 3, 4, 8))
def test_pull_openmp_equals_push_for_thread_counts(oracle_mod, threads):
    solid = walls_block_mask()
    oracle_mod.set_threads(threads)
    try:
        for prec in ("float", "double"):
            a = oracle_mod.Oracle(53, 31, 0.8, 1e-5, 2e-6, solid, prec, "v0")
            b = oracle_mod.Oracle(53, 31, 0.8, 1e-5, 2e-6, solid, prec, "pull")
            for o in (a, b):
                o.init_equilibrium(1.0, 0.02, 0.01)
                o.run(200)
            np.testing.assert_array_equal(bits(a.populations()), bits(b.populations()))
    finally:
        oracle_mod.set_threads(8)


def test_weights_are_correctly_rounded(oracle_mod):
    for i, w in enumerate([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4):
        assert oracle_mod.weight(i) == w
        assert oracle_mod.weight(i, "float") == float(np.float32(w))


CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def test_equilibrium_and_guo_moment_identities(oracle_mod):
    """sum g_eq = drho, sum c g_eq = rho u, sum c c g_eq = rho (u u) + (rho - 1) I / 3 ... (second
    moment of f_eq - w); sum guo = 0, sum c guo = F. Random inputs, FP64 to round-off."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        drho = rng.uniform(-0.05, 0.05)
        rho = 1.0 + drho
        ux, uy = rng.uniform(-0.1, 0.1, 2)
        fx, fy = rng.uniform(-1e-3, 1e-3, 2)
        usq = ux * ux + uy * uy
        w = np.array([oracle_mod.weight(i) for i in range(9)])
        geq = np.array([oracle_mod.equilibrium(w[i], drho, rho, CX[i] * ux + CY[i] * uy, usq)
                        for i in range(9)])
        src = np.array([oracle_mod.guo(w[i], CX[i], CY[i], ux, uy, fx, fy) for i in range(9)])
        assert abs(geq.sum() - drho) < 1e-15
        assert abs((CX * geq).sum() - rho * ux) < 1e-15
        assert abs((CY * geq).sum() - rho * uy) < 1e-15
        # second moments of f_eq = w + g_eq: rho (u u + I/3)
        f = w + geq
        assert abs((CX * CX * f).sum() - rho * (ux * ux + 1 / 3)) < 1e-14
        assert abs((CX * CY * f).sum() - rho * ux * uy) < 1e-14
        assert abs(src.sum()) < 1e-17
        assert abs((CX * src).sum() - fx) < 1e-17
        assert abs((CY * src).sum() - fy) < 1e-17


def test_collision_fixed_points(oracle_mod):
    # rest state without force stays exactly zero
    g = oracle_mod.collide(np.zeros(9), 0.0, 0.0, 1 / 0.8, 1 - 0.5 / 0.8)
    assert np.all(g == 0.0)
    # at tau = 1 (omega = 1): post-collision = g_eq(moments) + 1/2 guo
    rng = np.random.default_rng(5)
    g0 = rng.standard_normal(9) * 1e-3
    fx, fy = 2e-5, -1e-5
    out = oracle_mod.collide(g0, fx, fy, 1.0, 0.5)
    drho = g0.sum()
    rho = 1 + drho
    ux = ((CX * g0).sum() + 0.5 * fx) / rho
    uy = ((CY * g0).sum() + 0.5 * fy) / rho
    w = np.array([oracle_mod.weight(i) for i in range(9)])
    cu = CX * ux + CY * uy
    geq = w * (drho + rho * (3 * cu + 4.5 * cu * cu - 1.5 * (ux * ux + uy * uy)))
    src = w * (3 * ((CX - ux) * fx + (CY - uy) * fy) + 9 * cu * (CX * fx + CY * fy))
    np.testing.assert_allclose(out, geq + 0.5 * src, rtol=0, atol=1e-15)
    # mass conserved, momentum gains exactly F (in exact arithmetic): sum c (g' - g) = F
    for tau in (0.6, 0.8, 1.7):
        om = 1 / tau
        out = oracle_mod.collide(g0, fx, fy, om, 1 - 0.5 * om)
        assert abs(out.sum() - g0.sum()) < 1e-15
        assert abs((CX * out).sum() - (CX * g0).sum() - fx) < 1e-15
        assert abs((CY * out).sum() - (CY * g0).sum() - fy) < 1e-15


def test_periodic_box_momentum_grows_by_F_per_cell_per_step(oracle_mod):
    """No walls: streaming conserves momentum, each fluid cell gains F per step."""
    nx, ny, fx, fy = 12, 9, 3e-6, -2e-6
    o = oracle_mod.Oracle(nx, ny, 0.7, fx, fy, np.zeros(nx * ny, np.uint8), "double", "v0")
    o.init_equilibrium(1.0, 0.01, 0.02)
    p0 = o.momentum()
    o.run(50)
    p1 = o.momentum()
    assert abs(p1[0] - p0[0] - 50 * nx * ny * fx) < 1e-12
    assert abs(p1[1] - p0[1] - 50 * nx * ny * fy) < 1e-12
    assert abs(o.total_mass() - nx * ny) < 1e-12


def test_bounce_back_reflects_momentum_at_a_wall(oracle_mod):
    """A single population heading into a wall comes back reversed one step later, in the
    same cell (halfway bounce-back), and nothing leaks into the wall."""
    nx, ny = 5, 4
    solid = np.zeros((ny, nx), np.uint8)
    solid[0, :] = 1
    solid = solid.reshape(-1)
    S = np.zeros(9 * nx * ny)
    c = 1 * nx + 2  # cell (2, 1), next to the wall row y = 0
    S[c * 9 + 4] = 1e-3  # heading south (direction 4)
    o = oracle_mod.Oracle(nx, ny, 1e9, 0.0, 0.0, solid, "double", "v0")  # omega ~ 0: pure stream
    o.set_populations(S)
    o.run(1)
    out = o.populations().reshape(-1, 9)
    assert abs(out[c, 2] - 1e-3) < 1e-12  # came back north in the same cell
    assert np.abs(np.delete(out.reshape(-1), c * 9 + 2)).max() < 1e-9


def test_preimage_round_trip(oracle_mod):
    nx, ny = 53, 31
    solid = walls_block_mask()
    rng = np.random.default_rng(9)
    S = rng.standard_normal(9 * nx * ny)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    R = oracle_mod.aos_to_pull(S, nx, ny, solid)
    np.testing.assert_array_equal(oracle_mod.pull_to_aos(R, nx, ny, solid), S)


def _oracle_poiseuille(oracle_mod, H, tau, every=1000, steady_tol=1e-10):
    ny = H + 2
    force = 8.0 * ((tau - 0.5) / 3.0) * 0.05 / (H * H)
    nu = (tau - 0.5) / 3.0
    o = oracle_mod.Oracle(4, ny, tau, force, 0.0, oracle_mod.make_channel(4, ny), "double", "v0")
    o.init_equilibrium(1.0, 0.0, 0.0)
    m0 = o.total_mass()
    prev = None
    for _ in range(int(20 * H * H / nu / every)):
        o.run(every)
        cur = o.macroscopic()[1].reshape(ny, 4)[1:-1].mean(axis=1)
        if prev is not None and np.sqrt(np.sum((cur - prev) ** 2) / np.sum(prev ** 2)) < steady_tol:
            break
        prev = cur
    k = np.arange(H) + 0.5
    exact = force / (2 * nu) * k * (H - k)
    return math.sqrt(np.sum((cur - exact) ** 2) / np.sum(exact ** 2)), (o.total_mass() - m0) / m0


def test_oracle_poiseuille_exact_tau_and_order(oracle_mod):
    l2, drift = _oracle_poiseuille(oracle_mod, 32, 0.5 + math.sqrt(3) / 4)
    assert l2 < 1e-8 and abs(drift) < 1e-13
    e16, _ = _oracle_poiseuille(oracle_mod, 16, 1.0)
    e32, _ = _oracle_poiseuille(oracle_mod, 32, 1.0)
    assert 1.9 < math.log2(e16 / e32) < 2.1
