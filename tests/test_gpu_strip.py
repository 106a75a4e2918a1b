"""Strip engine: one GPU's share of the v4 strip decomposition (src/v4_mpi_cuda/), ghost rows as
kernel inputs. Reference test: tests/test_v4_vs_v0.cpp:53-71 (mpiexec -np 1..4, uneven strips,
overlap on and off, the walls + block channel and the open-x cylinder, each against v0). Here
the ring of strips is driven in one process on the one GPU, the host doing the exchange the MPI
ranks would do (solver_v4.cpp:105-130), and the owned rows are compared bit for bit with the
reference's own v0 output (the golden fixtures) and with the oracle, never with another path of
this library. One ghost row is the reference's strip (one step per exchange); two ghost rows
take two fused steps per exchange (compute_rows2)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def assert_bitwise(got, ref, what):
    diff = np.flatnonzero(bits(got) != bits(ref))
    assert diff.size == 0, (f"{what}: {diff.size} differing values, first at cell {diff[0] // 9} "
                            f"direction {diff[0] % 9}: {got[diff[0]]!r} vs {ref[diff[0]]!r}")


def strip_rows(y0, nyl, ny, ghost):
    """Global rows of a strip's local rows (ghost rows wrap around the ring)."""
    return [(y0 - ghost + ly) % ny for ly in range(nyl + 2 * ghost)]


def initial_pull_state(oracle_mod, nx, ny, tau, fx, fy, solid, prec, init, u_in=None):
    """The stored (pull) state of the reference's uniform initial equilibrium, from the oracle's
    pull restatement (pinned bit for bit to v0)."""
    o = oracle_mod.Oracle(nx, ny, tau, fx, fy, solid, prec, "pull")
    if u_in is not None:
        o.set_inflow_outflow(u_in)
    o.init_equilibrium(*init)
    return o.raw_state()


def boundary_then_interior(st, nyl, two, overlap, up=None, down=None):
    """SolverV4::step's launch order (solver_v4.cpp:119-123): the rows the neighbours need on
    stream 0, the interior on stream 1. A message of a strip with two ghost rows carries its
    two outermost owned rows on each side, so those are its boundary rows (for single steps
    too): pack_halo waits for stream 0 only. With `up` / `down` the two-step launches deliver
    the halo into those strips themselves (compute_rows2_to)."""
    g = st.ghost
    lo, hi = g, g + nyl - 1
    run = st.compute_rows2 if two else st.compute_rows
    if two and up is not None:
        def run(a, b, s):
            st.compute_rows2_to(a, b, s, up, down)
    depth = g
    if not overlap or nyl <= 2 * depth:
        run(lo, hi, 0)
        return
    run(lo, lo + depth - 1, 0)
    run(hi - depth + 1, hi, 0)
    run(lo + depth, hi - depth, 1)


def run_ring(lbm, nx, ny, heights, solid, R0, schedule, prec, tau, fx, fy, ghost,
             open_x=False, u_in=0.0, overlap=True, tile=None, deliver=False):
    """A ring of strips stepped by `schedule` (a list of 1 / 2: steps per exchange); returns the
    assembled pull state of the owned rows. deliver: two-step rounds store the halo into the
    neighbours' ghost rows from the kernel (no pack/unpack)."""
    assert sum(heights) == ny
    strips, starts = [], np.cumsum([0] + list(heights[:-1]))
    for y0, nyl in zip(starts, heights):
        rows = strip_rows(y0, nyl, ny, ghost)
        st = lbm.Strip(nx, nyl, prec, ghost)
        st.set_params(tau, fx, fy, open_x, u_in)
        if tile:
            st.set_two_step_tile(*tile)
        st.upload_solid(solid.reshape(ny, nx)[rows].reshape(-1))
        st.upload_state(np.ascontiguousarray(R0.reshape(9, ny, nx)[:, rows, :]).reshape(-1))
        strips.append((st, y0, nyl))
    K = len(strips)
    for n in schedule:
        fused = deliver and n == 2
        for k, (st, _, nyl) in enumerate(strips):
            if fused:
                boundary_then_interior(st, nyl, True, overlap, strips[(k + 1) % K][0],
                                       strips[(k - 1) % K][0])
            else:
                boundary_then_interior(st, nyl, n == 2, overlap)
        if not fused:
            halos = [st.pack_halo() for st, _, _ in strips]  # (send_up, send_down)
            for k, (st, _, _) in enumerate(strips):
                st.unpack_halo(halos[(k - 1) % K][0], halos[(k + 1) % K][1])
        for st, _, _ in strips:
            st.finish_step()
    out = np.empty((9, ny, nx), dtype=R0.dtype)
    for st, y0, nyl in strips:
        g = st.download_state().reshape(9, nyl + 2 * ghost, nx)
        out[:, y0:y0 + nyl, :] = g[:, ghost:ghost + nyl, :]
        st.close()
    return out.reshape(-1)


def schedule_for(steps, ghost, mixed=False):
    if ghost == 1:
        return [1] * steps
    if mixed:  # a single step, pairs, and a single step at the end
        return [1] + [2] * ((steps - 2) // 2) + [1] * (steps - 2 - 2 * ((steps - 2) // 2) + 1)
    return [2] * (steps // 2) + [1] * (steps % 2)


HEIGHTS = {1: [(31,), (16, 15), (10, 11, 10), (7, 8, 9, 7), (1, 29, 1)],
           2: [(31,), (16, 15), (10, 11, 10), (7, 8, 9, 7), (2, 27, 2), (3, 25, 3)]}


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("ghost,heights", [(g, h) for g in (1, 2) for h in HEIGHTS[g]])
def test_ring_of_strips_matches_reference_v0(lbm, oracle_mod, prec, ghost, heights):
    """53x31 walls + block, 500 steps, init (1, 0.02, 0.01) (test_v4_vs_v0.cpp:22-37): against
    the golden v0 output; overlap on for an even number of strips, off for an odd one."""
    print(f"strip ring {prec} ghost {ghost} heights {heights}", flush=True)
    solid, ref, p = load_golden(f"walls_block_53x31_500_{prec}.npz")
    nx, ny = p["nx"], p["ny"]
    R0 = initial_pull_state(oracle_mod, nx, ny, p["tau"], p["fx"], p["fy"], solid, prec, p["init"])
    R = run_ring(lbm, nx, ny, heights, solid, R0, schedule_for(p["steps"], ghost), prec,
                 p["tau"], p["fx"], p["fy"], ghost, overlap=len(heights) % 2 == 0)
    S = oracle_mod.pull_to_aos(R, nx, ny, solid)
    assert_bitwise(S, ref, f"{prec} ghost {ghost} strips {heights}")


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("heights", HEIGHTS[2])
def test_kernel_delivered_halo_matches_reference_v0(lbm, oracle_mod, prec, heights):
    """Two ghost rows, halo stored by each two-step launch straight into the neighbours' next
    ghost rows (strips of uneven height receive at different rows; one strip is its own
    neighbour), mixed with single steps exchanged through pack/unpack: golden v0 bits."""
    print(f"strip ring delivered halo {prec} heights {heights}", flush=True)
    solid, ref, p = load_golden(f"walls_block_53x31_500_{prec}.npz")
    nx, ny = p["nx"], p["ny"]
    R0 = initial_pull_state(oracle_mod, nx, ny, p["tau"], p["fx"], p["fy"], solid, prec, p["init"])
    R = run_ring(lbm, nx, ny, heights, solid, R0, schedule_for(p["steps"], 2, mixed=True), prec,
                 p["tau"], p["fx"], p["fy"], 2, overlap=len(heights) % 2 == 0, deliver=True)
    assert_bitwise(oracle_mod.pull_to_aos(R, nx, ny, solid), ref, f"{prec} delivered {heights}")


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("tile", [(64, 1, 0), (64, 3, 1), (128, 5, 2), (64, -1, 1)])
def test_two_step_strips_any_tile_and_mixed_steps(lbm, oracle_mod, prec, tile):
    """Two ghost rows: two-step tiles of any shape, and single steps mixed with pairs (one step
    first and last), same bits as v0 (golden)."""
    print(f"strip two-step tile {tile} {prec}", flush=True)
    solid, ref, p = load_golden(f"walls_block_53x31_500_{prec}.npz")
    nx, ny = p["nx"], p["ny"]
    R0 = initial_pull_state(oracle_mod, nx, ny, p["tau"], p["fx"], p["fy"], solid, prec, p["init"])
    sched = schedule_for(p["steps"], 2, mixed=True)
    assert sum(sched) == p["steps"] and sched[0] == 1 and sched[-1] == 1
    R = run_ring(lbm, nx, ny, (12, 9, 10), solid, R0, sched, prec, p["tau"], p["fx"], p["fy"], 2,
                 tile=tile)
    assert_bitwise(oracle_mod.pull_to_aos(R, nx, ny, solid), ref, f"{prec} tile {tile}")


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("ghost", (1, 2))
def test_ring_with_open_x_cylinder_matches_reference_v0(lbm, oracle_mod, prec, ghost):
    """The 120x48 open-x cylinder, 1000 steps (cylinder_case.hpp:16-31), uneven strips
    (strip_update.hpp:58-64: outlet source without y wrap inside a strip); single steps (the
    two-step walk takes periodic x only)."""
    print(f"strip ring cylinder {prec} ghost {ghost}", flush=True)
    solid, ref, p = load_golden(f"cylinder_120x48_1000_{prec}.npz")
    nx, ny = p["nx"], p["ny"]
    R0 = initial_pull_state(oracle_mod, nx, ny, p["tau"], 0.0, 0.0, solid, prec, p["init"],
                            u_in=p["u_in"])
    R = run_ring(lbm, nx, ny, (20, 13, 15), solid, R0, [1] * p["steps"], prec, p["tau"], 0.0,
                 0.0, ghost, True, p["u_in"])
    S = oracle_mod.pull_to_aos(R, nx, ny, solid, True, p["u_in"])
    assert_bitwise(S, ref, f"cylinder {prec} ghost {ghost}")


@pytest.mark.parametrize("ghost,prec", [(1, "float"), (2, "float"), (2, "double")])
def test_strip_at_real_shape_matches_oracle_channel(lbm, oracle_mod, ghost, prec):
    """8192 x (1024 + 2 ghost rows): one GPU's share out of 8 of the 8192^2 channel. With the
    ring closed on itself (device-side ghost refresh, the timed path) and walls on the strip's
    first and last rows it is a periodic-x walled channel of 1024 rows, whose flow is
    x-invariant: every column must equal the oracle's v0 run of a 4-column channel of the same
    height (tau 0.6, gx 1e-6, rest init), bit for bit."""
    print(f"strip real shape ghost {ghost} {prec}", flush=True)
    nx, nyl, steps = 8192, 1024, 20
    solid = lbm.make_channel(nx, nyl)
    rows = strip_rows(0, nyl, nyl, ghost)
    st = lbm.Strip(nx, nyl, prec, ghost)
    st.set_params(0.6, 1e-6, 0.0)
    st.upload_solid(solid.reshape(nyl, nx)[rows].reshape(-1))
    dt = np.float32 if prec == "float" else np.float64
    st.upload_state(np.zeros(9 * st.N, dt))  # rest: g = f - w = 0
    assert st.timed_steps(steps) > 0
    g = st.download_state().reshape(9, nyl + 2 * ghost, nx)[:, ghost:ghost + nyl, :]
    st.close()
    o = oracle_mod.Oracle(4, nyl, 0.6, 1e-6, 0.0, oracle_mod.make_channel(4, nyl), prec, "v0")
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.run(steps)
    R4 = oracle_mod.aos_to_pull(o.populations(), 4, nyl, oracle_mod.make_channel(4, nyl))
    col = R4.reshape(9, nyl, 4)[:, :, :1]
    assert np.abs(col).max() > 0  # the force has moved the flow
    mism = int(np.count_nonzero(bits(g) != bits(np.broadcast_to(col, g.shape))))
    assert mism == 0, f"{mism} values differ from the oracle channel"


def test_strip_rejects_bad_calls(lbm):
    st = lbm.Strip(64, 8, "float")
    with pytest.raises(ValueError):
        st.compute_rows(0, 3)
    with pytest.raises(ValueError):
        st.compute_rows(1, 9)
    with pytest.raises(ValueError, match="two ghost rows"):
        st.compute_rows2(1, 8)
    with pytest.raises(ValueError, match="halo messages"):
        st.unpack_halo(np.zeros(3 * 64 - 1, np.float32), np.zeros(3 * 64, np.float32))
    st.close()
    st = lbm.Strip(64, 8, "double", 2)
    assert st.halo_values == 12 * 64
    with pytest.raises(ValueError):
        st.compute_rows2(1, 8)  # row 1 is a ghost row
    with pytest.raises(ValueError):
        st.compute_rows(2, 10)
    st.set_params(0.8, 0.0, 0.0, True, 0.05)
    with pytest.raises(ValueError, match="periodic x"):
        st.compute_rows2(2, 9)
    other = lbm.Strip(64, 8, "double", 1)
    with pytest.raises(ValueError, match="two ghost rows"):
        st.compute_rows2_to(2, 9, 0, other, st)
    other.close()
    wide = lbm.Strip(128, 8, "double", 2)
    st.set_params(0.8)
    with pytest.raises(ValueError, match="same width"):
        st.compute_rows2_to(2, 9, 0, st, wide)
    wide.close()
    st.close()
    with pytest.raises(ValueError):
        lbm.Strip(64, 1, "float", 2)  # fewer owned rows than ghost rows
    with pytest.raises(ValueError):
        lbm.Strip(64, 8, "float", 3)
