"""Strip engine: one GPU's share of the v4 strip decomposition (src/v4_mpi_cuda/), ghost rows as
kernel inputs. Reference test: tests/test_v4_vs_v0.cpp (mpiexec -np 1..4, uneven strips); here
the ring of strips is driven in one process on the one GPU, with the host doing the exchange
the MPI ranks would do (solver_v4.cpp:105-130). Bar: owned rows bit-identical to the
un-split step."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import walls_block_mask

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def strip_rows(y0, nyl, ny):
    """Global rows of a strip's local rows 0..nyl+1 (ghost rows wrap around the ring)."""
    return [(y0 - 1 + ly) % ny for ly in range(nyl + 2)]


def split_state(R, nx, ny, rows):
    P = R.reshape(9, ny, nx)
    return np.ascontiguousarray(P[:, rows, :]).reshape(-1)


def run_ring(lbm, nx, ny, heights, solid, R0, steps, prec, tau, fx, fy, open_x=False, u_in=0.0,
             overlap=True):
    assert sum(heights) == ny
    strips, starts = [], np.cumsum([0] + list(heights[:-1]))
    for y0, nyl in zip(starts, heights):
        rows = strip_rows(y0, nyl, ny)
        st = lbm.Strip(nx, nyl, prec)
        st.set_params(tau, fx, fy, open_x, u_in)
        st.upload_solid(solid.reshape(ny, nx)[rows].reshape(-1))
        st.upload_state(split_state(R0, nx, ny, rows))
        strips.append((st, y0, nyl))
    K = len(strips)
    for _ in range(steps):
        for st, _, nyl in strips:
            if overlap:
                st.compute_rows(1, 1, 0)
                if nyl > 1:
                    st.compute_rows(nyl, nyl, 0)
                st.compute_rows(2, nyl - 1, 1)
            else:
                st.compute_rows(1, nyl, 0)
        halos = [st.pack_halo() for st, _, _ in strips]  # (send_up, send_down)
        for k, (st, _, _) in enumerate(strips):
            st.unpack_halo(halos[(k - 1) % K][0], halos[(k + 1) % K][1])
        for st, _, _ in strips:
            st.finish_step()
    out = np.empty((9, ny, nx), dtype=R0.dtype)
    for st, y0, nyl in strips:
        g = st.download_state().reshape(9, nyl + 2, nx)
        out[:, y0:y0 + nyl, :] = g[:, 1:nyl + 1, :]
    return out.reshape(-1)


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("heights", [(31,), (16, 15), (10, 11, 10), (7, 8, 9, 7), (1, 29, 1)])
def test_ring_of_strips_equals_unsplit_step(lbm, prec, heights):
    nx, ny, steps = 53, 31, 120
    solid = walls_block_mask()
    ref = lbm.Solver(nx, ny, 0.8, 1e-5, 2e-6, solid, prec)
    ref.init_equilibrium(1.0, 0.02, 0.01)
    R0 = ref.get_state()
    ref.run(steps)
    got = run_ring(lbm, nx, ny, heights, solid, R0, steps, prec, 0.8, 1e-5, 2e-6,
                   overlap=len(heights) % 2 == 0)
    Rref = ref.get_state().reshape(9, ny, nx)
    fluid = ~solid.reshape(ny, nx).astype(bool)
    a = got.reshape(9, ny, nx)[:, fluid]
    b = Rref[:, fluid]
    assert np.count_nonzero(bits(a) != bits(b)) == 0


@pytest.mark.parametrize("prec", ("float", "double"))
def test_ring_with_open_x_cylinder(lbm, prec):
    """Open-x inlet/outlet rules inside strips (strip_update.hpp:58-64: outlet source without y
    wrap inside a strip)."""
    nx, ny, steps = 120, 48, 200
    solid = lbm.make_cylinder(nx, ny, 30.0, 25.0, 10.0)
    tau = 0.575
    ref = lbm.Solver(nx, ny, tau, 0.0, 0.0, solid, prec)
    ref.set_inflow_outflow(0.05)
    ref.init_equilibrium(1.0, 0.05, 0.0)
    R0 = ref.get_state()
    ref.run(steps)
    got = run_ring(lbm, nx, ny, (20, 13, 15), solid, R0, steps, prec, tau, 0.0, 0.0, True, 0.05)
    fluid = ~solid.reshape(ny, nx).astype(bool)
    a = got.reshape(9, ny, nx)[:, fluid]
    b = ref.get_state().reshape(9, ny, nx)[:, fluid]
    assert np.count_nonzero(bits(a) != bits(b)) == 0


def test_strip_at_real_shape_self_ring(lbm):
    """8192 x (1024 + 2): the share of one GPU out of 8 for the 8192^2 channel. With the ring
    closed on itself (device-side ghost refresh) it is a periodic 8192 x 1024 domain."""
    nx, nyl, steps = 8192, 1024, 20
    solid = np.zeros((nyl, nx), np.uint8)
    solid[3, :] = 1  # a wall row inside the strip, and one on a ghost row
    solid[nyl - 1, : nx // 2] = 1
    solid = solid.reshape(-1)
    ref = lbm.Solver(nx, nyl, 0.6, 1e-6, 0.0, solid, "float")
    ref.init_equilibrium(1.0, 0.0, 0.0)
    R0 = ref.get_state()
    ref.run(steps)
    rows = strip_rows(0, nyl, nyl)
    st = lbm.Strip(nx, nyl, "float")
    st.set_params(0.6, 1e-6, 0.0)
    st.upload_solid(solid.reshape(nyl, nx)[rows].reshape(-1))
    st.upload_state(split_state(R0, nx, nyl, rows))
    sec = st.timed_steps(steps)
    assert sec > 0
    g = st.download_state().reshape(9, nyl + 2, nx)[:, 1:nyl + 1, :]
    fluid = ~solid.reshape(nyl, nx).astype(bool)
    assert np.count_nonzero(bits(g[:, fluid]) != bits(ref.get_state().reshape(9, nyl, nx)[:, fluid])) == 0


def test_strip_rejects_bad_rows(lbm):
    st = lbm.Strip(64, 8, "float")
    with pytest.raises(ValueError):
        st.compute_rows(0, 3)
    with pytest.raises(ValueError):
        st.compute_rows(1, 9)
