"""Discrete adjoint with binomial checkpointing, checked against central finite differences of an
FP64 forward run (the FP64 GPU step is itself bit-identical to the reference's v0 in double).

Objective: J = mean u_x over a probe window, averaged over the last `avg` states of a run from
rest, with a per-cell force field F and relaxation time tau. Gradients dJ/dF (per cell) and
dJ/dtau. Tolerances: double adjoint vs double FD 1e-6 relative; float adjoint vs double FD
1e-3 relative (the bar of the differentiation instance)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def disc_channel(lbm, nx, ny, xc, yc, r):
    return lbm.add_disc(lbm.make_channel(nx, ny), nx, ny, xc, yc, r)


def force_field(nx, ny, seed, gx=1e-5, sigma=1e-7):
    rng = np.random.default_rng(seed)
    F = rng.normal(0.0, sigma, (2, ny, nx))
    F[0] += gx
    return F


def fd_directional(lbm, geom, F, tau, v, h, steps, probe, avg):
    nx, ny, solid = geom
    a = lbm.Adjoint(nx, ny, tau, solid, steps, probe, avg, 2, "double")
    a.set_force_field(F + h * v)
    jp, _ = a.forward()
    a.set_force_field(F - h * v)
    jm, _ = a.forward()
    return (jp - jm) / (2 * h)


def fd_tau(lbm, geom, F, tau, dt, steps, probe, avg):
    nx, ny, solid = geom
    a = lbm.Adjoint(nx, ny, tau + dt, solid, steps, probe, avg, 2, "double")
    a.set_force_field(F)
    jp, _ = a.forward()
    a.set_tau(tau - dt)
    jm, _ = a.forward()
    return (jp - jm) / (2 * dt)


@pytest.mark.parametrize("prec,tol", (("double", 1e-6), ("float", 1e-3)))
def test_small_case_matches_finite_differences(lbm, prec, tol):
    nx, ny, steps, avg = 64, 32, 80, 20
    probe = (40, 10, 8, 12)
    solid = disc_channel(lbm, nx, ny, 20.0, 16.0, 4.0)
    geom = (nx, ny, solid)
    F = force_field(nx, ny, 1, gx=2e-5)
    tau = 0.7
    a = lbm.Adjoint(nx, ny, tau, solid, steps, probe, avg, 6, prec)
    a.set_force_field(F)
    J, dF, dtau, _, _ = a.gradient()
    Jf, _ = a.forward()
    assert abs(J - Jf) <= 1e-12 * abs(Jf) + 1e-18
    rng = np.random.default_rng(2)
    fluid = ~solid.reshape(ny, nx).astype(bool)
    for _ in range(4):
        v = rng.standard_normal((2, ny, nx)) * fluid
        fd = fd_directional(lbm, geom, F, tau, v, 1e-6, steps, probe, avg)
        ad = float(np.sum(dF.astype(np.float64) * v))
        assert abs(ad - fd) <= tol * abs(fd), (ad, fd)
    fdt = fd_tau(lbm, geom, F, tau, 1e-5, steps, probe, avg)
    assert abs(dtau - fdt) <= tol * abs(fdt), (dtau, fdt)


@pytest.mark.parametrize("prec", ("float", "double"))
def test_checkpoint_schedules_give_identical_gradients(lbm, prec):
    """Recomputed states are bit-identical to stored ones, so every schedule and snapshot budget
    gives the same bits: binomial schedules (their recomputation count follows the binomial
    optimum) and pipelined ones (one recomputation step per level inside each reverse launch)."""
    nx, ny, steps, avg = 64, 24, 37, 9
    solid = disc_channel(lbm, nx, ny, 30.0, 12.0, 3.0)
    F = force_field(nx, ny, 3)
    ref = None
    counts = {}
    runs = [("binomial", n) for n in (1, 2, 3, 6, 38, 100)] + \
           [("pipelined", n) for n in (9, 10, 14, 38)]
    for kind, snaps in runs:
        a = lbm.Adjoint(nx, ny, 0.8, solid, steps, (44, 6, 6, 10), avg, snaps, prec, kind)
        info = a.schedule_info()
        assert info["kind"] == kind and info["slots"] <= snaps
        a.set_force_field(F)
        J, dF, dtau, _, fsteps = a.gradient()
        a.close()
        if kind == "binomial":
            counts[snaps] = fsteps
        else:
            assert info["launches"] == 2 * steps + 1  # every reverse launch carries the chains
        if ref is None:
            ref = (J, dF.copy(), dtau)
        else:
            assert np.array_equal(dF, ref[1]), (kind, snaps)
            assert dtau == ref[2], (kind, snaps)
            assert J == ref[0], (kind, snaps)
    n = steps + 1
    assert counts[1] == n * (n - 1) // 2   # no free slot: recompute from R_0 every time
    assert counts[38] == steps and counts[100] == steps  # every state stored once
    assert counts[1] > counts[2] > counts[3] > counts[6] > counts[38]


@pytest.mark.parametrize("prec", ("float", "double"))
def test_forward_matches_the_plain_step(lbm, prec):
    """The adjoint's forward steps are the plain step's arithmetic (collide() on the same pulled
    values): J from the adjoint's forward run equals J from the plain solver (bit-identical to
    the reference's v0) stepped with the same force field, read back state by state."""
    nx, ny, steps, avg = 48, 20, 30, 7
    probe = (30, 5, 6, 9)
    solid = disc_channel(lbm, nx, ny, 14.0, 10.0, 3.5)
    F = force_field(nx, ny, 9, gx=3e-5, sigma=1e-6)
    a = lbm.Adjoint(nx, ny, 0.75, solid, steps, probe, avg, 4, prec)
    a.set_force_field(F)
    J, _ = a.forward()
    Jg = a.gradient()[0]
    s = lbm.Solver(nx, ny, 0.75, 0.0, 0.0, solid, prec)
    s.set_force_field(F)
    s.init_equilibrium(1.0, 0.0, 0.0)
    x0, y0, w, h = probe
    acc = 0.0
    for t in range(steps + 1):
        if t > steps - avg:
            ux = s.macroscopic()[1].reshape(ny, nx)[y0:y0 + h, x0:x0 + w].astype(np.float64)
            acc += float(np.sum(ux))
        if t < steps:
            s.step()
    s.close()
    Jp = acc / (avg * w * h)
    assert abs(J - Jp) <= 1e-12 * abs(Jp), (J, Jp)
    assert abs(Jg - Jp) <= 1e-12 * abs(Jp), (Jg, Jp)


def test_baseline_instance_512x128_T2000(lbm):
    """512 x 128 channel, disc r = 12 at (128, 64), T = 2000, probe 16 x 32 at x = 384, J
    averaged over the last 200 states; F = N(0, 1e-7) + gx 1e-5 (seeded), tau = 0.7;
    checkpoints ~ sqrt(T) = 45. Float adjoint vs central FD of the FP64 forward along 16 seeded
    directions and along tau: 1e-3 relative."""
    nx, ny, steps, avg = 512, 128, 2000, 200
    probe = (384, 48, 16, 32)
    solid = disc_channel(lbm, nx, ny, 128.0, 64.0, 12.0)
    geom = (nx, ny, solid)
    F = force_field(nx, ny, 2024)
    tau = 0.7
    a = lbm.Adjoint(nx, ny, tau, solid, steps, probe, avg, 45, "float")
    info = a.schedule_info()
    assert info["kind"] == "pipelined" and info["slots"] <= 45
    a.set_force_field(F)
    J, dF, dtau, t_grad, fsteps = a.gradient()
    Jf, t_fwd = a.forward()
    assert abs(J - Jf) <= 1e-12 * abs(Jf)
    print(f"J={J:.6e} forward {t_fwd * 1e3:.2f} ms, gradient {t_grad * 1e3:.2f} ms "
          f"(x{t_grad / t_fwd:.2f}), forward steps {fsteps}, schedule {info}")
    rng = np.random.default_rng(77)
    fluid = ~solid.reshape(ny, nx).astype(bool)
    worst = 0.0
    for _ in range(16):
        v = rng.standard_normal((2, ny, nx)) * fluid
        fd = fd_directional(lbm, geom, F, tau, v, 1e-8, steps, probe, avg)
        ad = float(np.sum(dF.astype(np.float64) * v))
        worst = max(worst, abs(ad - fd) / abs(fd))
    fdt = fd_tau(lbm, geom, F, tau, 1e-5, steps, probe, avg)
    worst_tau = abs(dtau - fdt) / abs(fdt)
    print(f"worst relative error: force directions {worst:.2e}, tau {worst_tau:.2e}")
    assert worst <= 1e-3 and worst_tau <= 1e-3
