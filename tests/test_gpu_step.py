"""GPU parity of the fused step through the C ABI (reference tests/test_v3_vs_v0.cu,
tests/test_cylinder_v3.cu, tests/poiseuille_check.hpp).

Bar: float and double are bit-identical to the reference's v0 (golden fixtures written by the
reference itself, and the oracle restatement at the same seeded inputs). Comparisons use the raw
bit patterns (stricter than the reference's `!=`, which lets +0 and -0 pass)."""
from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, walls_block_mask

pytestmark = pytest.mark.gpu

PRECISIONS = ("float", "double")
SHAPES = [(64, 1), (64, 3), (128, 8), (192, 5), (256, 8), (512, 2), (1024, 1)]


def bits(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:  # raw FP16 storage
        return a
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def assert_bitwise(got, ref, what=""):
    diff = np.flatnonzero(bits(got) != bits(ref))
    assert diff.size == 0, (f"{what}: {diff.size} differing values, first at cell {diff[0] // 9} "
                            f"direction {diff[0] % 9}: {got[diff[0]]!r} vs {ref[diff[0]]!r}")


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("shape", SHAPES)
def test_walls_block_matches_reference_v0(lbm, prec, shape):
    """53x31, 500 steps, walls + block, init (1, 0.02, 0.01) (test_v3_vs_v0.cu:20-37)."""
    solid, ref, p = load_golden(f"walls_block_53x31_500_{prec}.npz")
    s = lbm.Solver(53, 31, p["tau"], p["fx"], p["fy"], solid, prec)
    s.set_block(*shape)
    s.init_equilibrium(*p["init"])
    s.run(p["steps"])
    assert_bitwise(s.populations(), ref, f"{prec} block {shape}")


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("shape", [(64, 1), (128, 4), (256, 8)])
def test_cylinder_open_x_matches_reference_v0(lbm, prec, shape):
    """120x48 cylinder D = 10, inlet/outlet, 1000 steps (tests/cylinder_case.hpp:16-31)."""
    solid, ref, p = load_golden(f"cylinder_120x48_1000_{prec}.npz")
    s = lbm.Solver(p["nx"], p["ny"], p["tau"], 0.0, 0.0, solid, prec)
    s.set_block(*shape)
    s.set_inflow_outflow(p["u_in"])
    s.init_equilibrium(*p["init"])
    s.run(p["steps"])
    assert_bitwise(s.populations(), ref, f"cylinder {prec} {shape}")


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("seed", (0, 1, 2))
def test_seeded_block_rest_init_matches_reference_v0(lbm, prec, seed):
    """53x31, rest init, 8x8 block placed by seed (validation instance of the plain method)."""
    solid, ref, p = load_golden(f"seeded_block_53x31_500_seed{seed}_{prec}.npz")
    s = lbm.Solver(53, 31, p["tau"], p["fx"], p["fy"], solid, prec)
    s.init_equilibrium(1.0, 0.0, 0.0)
    s.run(p["steps"])
    assert_bitwise(s.populations(), ref, f"seed {seed} {prec}")


def test_crop256_double_matches_reference_hash(lbm):
    """256x256 walled channel, tau 0.6, gx 1e-6, 200 steps, double: SHA-256 of the reference
    v0 output (the fp64 crop of the 8192^2 instance)."""
    import hashlib

    h = json.loads((GOLDEN / "hashes.json").read_text())["crop256_f64"]
    s = lbm.Solver(256, 256, h["tau"], h["fx"], h["fy"], lbm.make_channel(256, 256), "double")
    s.init_equilibrium(1.0, 0.0, 0.0)
    s.run(h["steps"])
    S = s.populations()
    np.testing.assert_array_equal(S[h["sample_index"]], np.array(h["sample_value"]))
    assert hashlib.sha256(S.tobytes()).hexdigest() == h["sha256"]
    assert abs(s.total_mass() - h["total_mass"]) <= 1e-12 * h["total_mass"]


@pytest.mark.parametrize("prec", PRECISIONS)
def test_macroscopic_and_reductions_match_oracle(lbm, oracle_mod, prec):
    solid = walls_block_mask()
    s = lbm.Solver(53, 31, 0.8, 1e-5, 2e-6, solid, prec)
    o = oracle_mod.Oracle(53, 31, 0.8, 1e-5, 2e-6, solid, prec, "v0")
    for x in (s, o):
        x.init_equilibrium(1.0, 0.02, 0.01)
        x.run(137)
    for a, b in zip(s.macroscopic(), o.macroscopic()):
        assert_bitwise(a, b, "macroscopic")
    m_gpu, m_ref = s.total_mass(), o.total_mass()
    assert abs(m_gpu - m_ref) <= 1e-12 * m_ref
    (px, py), (qx, qy) = s.momentum(), o.momentum()
    assert abs(px - qx) <= 1e-12 * (abs(qx) + 1e-300) + 1e-15
    assert abs(py - qy) <= 1e-12 * (abs(qy) + 1e-300) + 1e-15
    S = s.populations().reshape(-1, 9)
    for (x, y) in ((0, 1), (5, 9), (52, 29), (20, 15), (12, 10)):
        assert_bitwise(s.populations_at(x, y), S[y * 53 + x], f"populations_at {x},{y}")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_set_populations_round_trip(lbm, prec):
    """Upload an arbitrary post-streaming state: its preimage gathers back to the same values."""
    solid = walls_block_mask()
    rng = np.random.default_rng(7)
    dt = np.float32 if prec == "float" else np.float64
    S = (rng.standard_normal(53 * 31 * 9) * 1e-3).astype(dt)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    s = lbm.Solver(53, 31, 0.8, 1e-5, 2e-6, solid, prec)
    s.set_populations(S)
    assert_bitwise(s.populations(), S, "round trip")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_random_state_and_force_field_match_oracle(lbm, oracle_mod, prec):
    """Arbitrary initial state and a per-cell force field (extension used by the adjoint
    form), 200 steps, against the oracle with the same field."""
    solid = walls_block_mask()
    rng = np.random.default_rng(11)
    dt = np.float32 if prec == "float" else np.float64
    S = (rng.standard_normal(53 * 31 * 9) * 1e-3).astype(dt)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    F = (rng.standard_normal(2 * 53 * 31) * 1e-5).astype(dt)
    s = lbm.Solver(53, 31, 0.7, 0.0, 0.0, solid, prec)
    o = oracle_mod.Oracle(53, 31, 0.7, 0.0, 0.0, solid, prec, "v0")
    for x in (s, o):
        x.set_populations(S)
        x.set_force_field(F)
        x.run(200)
    assert_bitwise(s.populations(), o.populations(), "force field")
    for a, b in zip(s.macroscopic(), o.macroscopic()):
        assert_bitwise(a, b, "macroscopic with field")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_uniform_field_equals_uniform_force(lbm, prec):
    solid = walls_block_mask()
    N = 53 * 31
    dt = np.float32 if prec == "float" else np.float64
    F = np.concatenate([np.full(N, 1e-5), np.full(N, 2e-6)]).astype(dt)
    a = lbm.Solver(53, 31, 0.8, 1e-5, 2e-6, solid, prec)
    b = lbm.Solver(53, 31, 0.8, 0.0, 0.0, solid, prec)
    b.set_force_field(F)
    for x in (a, b):
        x.init_equilibrium(1.0, 0.02, 0.01)
        x.run(300)
    assert_bitwise(a.populations(), b.populations(), "uniform field")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_setters_match_oracle_restart(lbm, oracle_mod, prec):
    """set_tau / set_force / set_mask between runs == an oracle restarted from the state."""
    solid = walls_block_mask()
    s = lbm.Solver(53, 31, 0.8, 1e-5, 2e-6, solid, prec)
    o = oracle_mod.Oracle(53, 31, 0.8, 1e-5, 2e-6, solid, prec, "v0")
    for x in (s, o):
        x.init_equilibrium(1.0, 0.02, 0.01)
        x.run(100)
    S = o.populations()
    solid2 = solid.copy().reshape(31, 53)
    solid2[20:24, 30:34] = 1
    solid2 = solid2.reshape(-1)
    S.reshape(-1, 9)[solid2.astype(bool)] = 0
    s.set_populations(S)
    s.set_tau(0.65)
    s.set_force(-3e-6, 1e-6)
    s.set_mask(solid2)
    s.set_populations(S)
    s.run(150)
    o2 = oracle_mod.Oracle(53, 31, 0.65, -3e-6, 1e-6, solid2, prec, "v0")
    o2.set_populations(S)
    o2.run(150)
    assert_bitwise(s.populations(), o2.populations(), "setters")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_channel_8192_columns_match_oracle_narrow_channel(lbm, oracle_mod, prec):
    """Full-size instance, 8192 x 8192 walled channel, tau 0.6, gx 1e-6, 200 steps (float, and
    double: the fp64 grid of the same instance): the flow is x-invariant, so every column must
    equal, bit for bit, the oracle's 4 x 8192 channel with the same rows. A size-independent
    property checked at the real size."""
    nx = ny = 8192
    s = lbm.Solver(nx, ny, 0.6, 1e-6, 0.0, lbm.make_channel(nx, ny), prec)
    s.init_equilibrium(1.0, 0.0, 0.0)
    s.run(200)
    rho, ux, uy = (a.reshape(ny, nx) for a in s.macroscopic())
    o = oracle_mod.Oracle(4, ny, 0.6, 1e-6, 0.0, oracle_mod.make_channel(4, ny), prec, "v0")
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.run(200)
    r0, u0, v0 = (a.reshape(ny, 4)[:, :1] for a in o.macroscopic())
    for got, ref, name in ((rho, r0, "rho"), (ux, u0, "ux"), (uy, v0, "uy")):
        bad = np.count_nonzero(bits(got) != bits(np.broadcast_to(ref, got.shape)))
        assert bad == 0, f"{name}: {bad} cells differ"
    assert ux[ny // 2, 0] > 0
    m0 = float(nx) * (ny - 2)
    assert abs(s.total_mass() - m0) / m0 < 1e-6
    s.close()


@pytest.mark.parametrize("prec", PRECISIONS)
def test_block_2048_1000_steps_match_oracle(lbm, oracle_mod, prec):
    """The plain method's throughput instance at its real size: 2048 x 2048, walls at y = 0 and
    ny - 1, the 8 x 8 block placed by seed 0, tau 0.8, F = (1e-5, 2e-6), rest init, 1000 steps
    on the default path (two steps per launch), every population bit for bit against the
    reference's own v1 OpenMP solver compiled from its sources (oracle/_ref), or, where that was
    not built, the oracle's OpenMP pull port (bit-identical to the v0 restatement,
    tests/test_oracle.py)."""
    n, steps = 2048, 1000
    solid = lbm.seeded_block_mask(0, n, n)
    s = lbm.Solver(n, n, 0.8, 1e-5, 2e-6, solid, prec)
    assert s.steps_per_launch()["in_use"]
    s.init_equilibrium(1.0, 0.0, 0.0)
    s.run(steps)
    got = s.populations()
    s.close()
    if oracle_mod.Reference.available(prec):
        ref = oracle_mod.Reference(prec).run("v1", n, n, 0.8, 1e-5, 2e-6, solid, steps)
    else:
        o = oracle_mod.Oracle(n, n, 0.8, 1e-5, 2e-6, solid, prec, "pull")
        o.init_equilibrium(1.0, 0.0, 0.0)
        o.run(steps)
        ref = o.populations()
    assert_bitwise(got, ref, f"2048^2 block {prec}")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_non_multiple_grid_and_tiny_grid(lbm, oracle_mod, prec):
    """Ragged sizes: nx not a multiple of 64, tiles cut in y, a 1-row-thick fluid layer."""
    for nx, ny in ((67, 9), (130, 5), (3, 3), (1, 5)):
        solid = np.zeros(nx * ny, np.uint8)
        solid[:nx] = 1
        s = lbm.Solver(nx, ny, 0.9, 2e-5, -1e-5, solid, prec)
        s.set_block(64, 4)
        o = oracle_mod.Oracle(nx, ny, 0.9, 2e-5, -1e-5, solid, prec, "v0")
        for x in (s, o):
            x.init_equilibrium(1.0, 0.01, -0.02)
            x.run(37)
        assert_bitwise(s.populations(), o.populations(), f"{nx}x{ny}")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_padded_planes_match_oracle(lbm, oracle_mod, prec):
    """Planes of 1 MiB and more sit a padded stride apart in device memory (a layout choice of
    the library). 640 x 416 (266,240 cells): walls, a block, an arbitrary state, a per-cell
    force field, 30 steps, bit for bit against the oracle; the raw pull state survives a
    download/upload round trip; fields and momentum exchange read the padded planes."""
    nx, ny = 640, 416
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[150:190, 200:230] = 1
    solid = solid.reshape(-1)
    rng = np.random.default_rng(21)
    dt = np.float32 if prec == "float" else np.float64
    S = (rng.standard_normal(nx * ny * 9) * 1e-3).astype(dt)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    F = (rng.standard_normal(2 * nx * ny) * 1e-5).astype(dt)
    s = lbm.Solver(nx, ny, 0.7, 0.0, 0.0, solid, prec)
    info = json.loads("{" + s.bench_info() + "}")
    assert info["plane_stride"] > nx * ny
    o = oracle_mod.Oracle(nx, ny, 0.7, 0.0, 0.0, solid, prec, "v0")
    for x in (s, o):
        x.set_populations(S)
        x.set_force_field(F)
        x.run(30)
    assert_bitwise(s.populations(), o.populations(), "padded planes")
    for a, b in zip(s.macroscopic(), o.macroscopic()):
        assert_bitwise(a, b, "padded macroscopic")
    R = s.get_state()
    assert R.size == 9 * nx * ny
    s.set_state(R)
    assert_bitwise(s.get_state(), R, "raw state round trip")
    s.run(5)
    o.run(5)
    assert_bitwise(s.populations(), o.populations(), "after raw round trip")
    fx, fy = s.momentum_exchange(190, 140, 50, 60)
    assert np.isfinite(fx) and np.isfinite(fy) and (fx, fy) != (0.0, 0.0)


@pytest.mark.parametrize("prec", ("float", "double", "fp16"))
@pytest.mark.parametrize("order", (0, 1, 2))
def test_tile_orders_give_identical_bits(lbm, oracle_mod, prec, order):
    """The workgroup-to-tile order (round-robin over the XCDs, XCD bands, XCD rows) changes no
    value: 192 x 64 walls + block, 120 steps, against the oracle (float, double) or against the
    round-robin run (FP16 storage, whose stochastic rounding is seeded per cell and step)."""
    nx, ny = 192, 64
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[20:30, 40:52] = 1
    solid = solid.reshape(-1)
    runs = []
    for o in (0, order):
        s = lbm.Solver(nx, ny, 0.8, 1e-5, 2e-6, solid, prec)
        s.set_block(64, 1)
        nt, one, _ = s.kernel_options()
        s.set_kernel_options(nt, one, o)
        s.init_equilibrium(1.0, 0.02, 0.01)
        s.run(120)
        runs.append(s.populations())
        s.close()
    assert_bitwise(runs[1], runs[0], f"tile order {order}")
    if prec != "fp16":
        ref = oracle_mod.Oracle(nx, ny, 0.8, 1e-5, 2e-6, solid, prec, "v0")
        ref.init_equilibrium(1.0, 0.02, 0.01)
        ref.run(120)
        assert_bitwise(runs[1], ref.populations(), f"tile order {order} vs oracle")


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("grid,tile", [((53, 31), (512, 64)), ((53, 31), (64, 1)),
                                       ((200, 70), (64, 7)), ((640, 416), (512, 64)),
                                       ((640, 416), (128, 3)), ((1000, 300), (256, 16)),
                                       # tiles taller than one halo chunk (16 rows): several chunks,
                                       # a partial last one, rows past the four-row trips
                                       ((640, 416), (512, 150)), ((333, 700), (128, 333)),
                                       # the automatic height
                                       ((300, 257), (256, -1))])
def test_two_step_launches_match_oracle(lbm, oracle_mod, prec, grid, tile):
    """Temporal blocking: run(n) advances pairs of steps in one launch (tiles of block_x
    columns by rows rows, halo columns and rows recomputed), odd n ends with a single step.
    Walls, an obstacle touching a tile edge, an arbitrary state and a uniform force; bit for bit
    against the oracle after 1, 2, 5 and 8 steps."""
    nx, ny = grid
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[ny // 3: ny // 3 + 9, min(nx - 12, 60): min(nx - 12, 60) + 7] = 1
    solid[ny // 2, :: 7] = 1
    solid = solid.reshape(-1)
    rng = np.random.default_rng(nx * 7 + ny)
    dt = np.float32 if prec == "float" else np.float64
    S = (rng.standard_normal(nx * ny * 9) * 1e-3).astype(dt)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    if prec == "double":  # its LDS ring holds twice the bytes: at most 256 columns
        tile = (min(tile[0], 256), tile[1])
    s = lbm.Solver(nx, ny, 0.73, 2e-5, -7e-6, solid, prec)
    s.set_steps_per_launch(2, *tile)
    info = s.steps_per_launch()
    assert info["in_use"]
    if tile[1] > 0:
        assert info["rows"] == tile[1]
    else:
        assert info["rows"] >= 16
    o = oracle_mod.Oracle(nx, ny, 0.73, 2e-5, -7e-6, solid, prec, "v0")
    for x in (s, o):
        x.set_populations(S)
    done = 0
    for n in (1, 1, 3, 3):
        s.run(n)
        o.run(n)
        done += n
        assert_bitwise(s.populations(), o.populations(), f"{grid} tile {tile} after {done} steps")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_two_step_raw_state_matches_single_steps(lbm, prec):
    """The raw device state (pull form, solid cells included) after two-step launches equals
    the single-step path's bit for bit: solid cells keep the values the upload gave them (their
    results go to the discard slots past each plane), fluid cells match."""
    nx, ny = 300, 90
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[30:45, 100:130] = 1
    solid = solid.reshape(-1)
    rng = np.random.default_rng(11)
    dt = np.float32 if prec == "float" else np.float64
    R = (rng.standard_normal(nx * ny * 9) * 1e-2).astype(dt)  # solid cells hold values too
    out = []
    for spl in (2, 1):
        s = lbm.Solver(nx, ny, 0.7, 1e-5, 3e-6, solid, prec)
        s.set_steps_per_launch(spl, 256 if prec == "double" else 512, 40)
        assert s.steps_per_launch()["in_use"] == (spl == 2)
        s.set_state(R)
        s.run(6)
        out.append(s.get_state())
        s.close()
    assert_bitwise(out[0], out[1], "two-step vs single-step raw state")
    cells = np.asarray(R).reshape(9, ny * nx)[:, solid.astype(bool)]
    assert_bitwise(out[0].reshape(9, ny * nx)[:, solid.astype(bool)], cells, "solid cells kept")


def test_two_step_path_is_the_default_where_it_applies(lbm):
    s = lbm.Solver(256, 64, 0.8, 1e-5, 0.0, lbm.make_channel(256, 64), "float")
    assert s.steps_per_launch()["in_use"]
    s.set_inflow_outflow(0.01)
    assert s.steps_per_launch()["in_use"]   # open x: two steps too
    s2 = lbm.Solver(256, 64, 0.8, 0.0, 0.0, lbm.make_channel(256, 64), "fp16")
    assert s2.steps_per_launch()["in_use"]  # FP16 storage too
    s2.set_inflow_outflow(0.01)
    assert not s2.steps_per_launch()["in_use"]  # but FP16 with open x: single steps
    s3 = lbm.Solver(256, 64, 0.8, 0.0, 0.0, lbm.make_channel(256, 64), "double")
    s3.set_force_field(np.zeros(2 * 256 * 64))
    assert s3.steps_per_launch()["in_use"]  # per-cell force field: two steps


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("nx,block,field", [(513, 512, False), (130, 64, True), (513, 256, True),
                                             (200, 128, False)])
def test_two_step_open_x_and_force_field_match_single_steps(lbm, prec, nx, block, field):
    """Open x (inlet column 0, outlet column nx - 1) and a per-cell force field on the two-step
    walk: tiles whose last one holds the outlet column alone (nx = 513 with 512- or 256-column
    tiles: the outlet is a halo column of the tile before it), obstacles touching the inlet and
    the outlet (outlet bounce-back bits), walls; 9 steps (4 two-step launches and one single
    step) bit for bit equal to 9 single steps."""
    ny = 70
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[20:30, 0:6] = 1              # on the inlet
    solid[40:52, nx - 4:nx] = 1        # on the outlet
    solid[10:18, nx // 2:nx // 2 + 9] = 1
    solid = solid.reshape(-1)
    dt = np.float32 if prec == "float" else np.float64
    if prec == "double":  # its LDS ring holds twice the bytes: at most 256 columns
        block = min(block, 256)
    rng = np.random.default_rng(nx + block)
    F = (rng.standard_normal(2 * nx * ny) * 1e-5).astype(dt)
    out = []
    for spl in (2, 1):
        s = lbm.Solver(nx, ny, 0.7, 2e-6, 0.0, solid, prec)
        s.set_inflow_outflow(0.03)
        if field:
            s.set_force_field(F)
        s.set_steps_per_launch(spl, block, 16)
        assert s.steps_per_launch()["in_use"] == (spl == 2)
        s.init_equilibrium(1.0, 0.03, 0.0)
        s.run(9)
        out.append(s.populations())
        s.close()
    assert_bitwise(out[0], out[1], f"open x {prec} nx={nx} block={block} field={field}")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_autotune_restores_state_and_keeps_results(lbm, prec):
    """lbm_autotune times every one-step shape and option and, where run() takes it, every
    two-step tile; the state is restored and the tuned solver computes the untuned one's bits."""
    nx, ny = 640, 200
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[80:100, 300:330] = 1
    solid = solid.reshape(-1)
    rng = np.random.default_rng(5)
    S = (rng.standard_normal(nx * ny * 9) * 1e-3).astype(
        {"float": np.float32, "double": np.float64, "fp16": np.float32}[prec])
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    out = []
    for tune in (True, False):
        s = lbm.Solver(nx, ny, 0.7, 1e-5, 0.0, solid, prec)
        s.set_populations(S)
        if tune:
            before = s.populations()
            bx, r, mlups = s.autotune(steps=2)
            assert mlups > 0 and bx % 64 == 0 and r >= 1
            assert_bitwise(s.populations(), before, "state restored after autotune")
            info = s.steps_per_launch()
            assert info["in_use"] == (prec != "fp16")
        s.run(7)
        out.append(s.populations())
        s.close()
    assert_bitwise(out[0], out[1], "tuned vs untuned")


@pytest.mark.parametrize("prec", PRECISIONS)
def test_two_step_tile_orders_match_oracle(lbm, oracle_mod, prec):
    """The two-step kernel's tile orders (round-robin, XCD bands, XCD rows: 8 tile rows here)
    give the oracle's bits."""
    nx, ny = 640, 512
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[200:230, 500:520] = 1
    solid = solid.reshape(-1)
    o = oracle_mod.Oracle(nx, ny, 0.9, 3e-6, 1e-6, solid, prec, "v0")
    o.init_equilibrium(1.0, 0.03, 0.0)
    o.run(6)
    ref = o.populations()
    for order in (0, 1, 2):
        s = lbm.Solver(nx, ny, 0.9, 3e-6, 1e-6, solid, prec)
        s.set_steps_per_launch(2, 256, 64, order)
        info = s.steps_per_launch()
        assert info["in_use"] and info["tile_order"] == order
        s.init_equilibrium(1.0, 0.03, 0.0)
        s.run(6)
        assert_bitwise(s.populations(), ref, f"two-step tile order {order}")
        s.close()


@pytest.mark.parametrize("grid,tile", [((53, 31), (64, 7)), ((300, 90), (512, -1)),
                                       ((640, 416), (256, 64)), ((333, 700), (128, 100)),
                                       ((640, 512), (256, 64))])
def test_fp16_two_step_equals_one_step_launches(lbm, grid, tile):
    """FP16 storage with two steps per launch: each step's stochastic rounding (the hash of the
    cell and the step pair, read rotated on the odd step) applied to the step-1 values before
    step 2 reads them, so the raw state equals the one-step kernel's bit for bit after even and
    odd step counts (run(3) from step 3 takes one single step, then a two-step launch from an
    even step); solid cells keep their values."""
    nx, ny = grid
    solid = lbm.make_channel(nx, ny).reshape(ny, nx)
    solid[ny // 3: ny // 3 + 9, min(nx - 12, 40): min(nx - 12, 40) + 7] = 1
    solid = solid.reshape(-1)
    rng = np.random.default_rng(nx + 3 * ny)
    S = (rng.standard_normal(nx * ny * 9) * 2e-3).astype(np.float32)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    out = []
    for spl in (2, 1):
        s = lbm.Solver(nx, ny, 0.7, 2e-5, -5e-6, solid, "fp16")
        s.set_steps_per_launch(spl, *tile)
        assert s.steps_per_launch()["in_use"] == (spl == 2)
        assert s.steps_per_launch()["tile_order"] == 2  # the FP16 default: XCD rows (8 tile rows
        # at 640 x 512; the other grids fall back to round-robin)
        s.set_populations(S)
        states = []
        for n in (2, 1, 3, 6):
            s.run(n)
            states.append(s.get_state())
        out.append(states)
        s.close()
    for k, (a, b) in enumerate(zip(*out)):
        assert_bitwise(a, b, f"FP16 {grid} tile {tile} after run {k}")


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("grid", [(1, 1), (1, 7), (5, 1), (3, 3), (64, 2), (513, 20), (130, 5)])
def test_two_step_degenerate_grids_match_oracle(lbm, oracle_mod, prec, grid):
    """Fully periodic domains (no walls, an obstacle cell where there is room) down to one column
    or one row, and a last column tile one cell wide (513 = 512 + 1): the two-step path against
    the oracle, bit for bit, after odd and even step counts."""
    nx, ny = grid
    solid = np.zeros(nx * ny, np.uint8)
    if nx >= 3 and ny >= 3:
        solid[(ny // 2) * nx + nx // 2] = 1
    rng = np.random.default_rng(nx * 131 + ny)
    dt = np.float32 if prec == "float" else np.float64
    S = (rng.standard_normal(nx * ny * 9) * 1e-3).astype(dt)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    s = lbm.Solver(nx, ny, 0.9, 1e-5, -2e-6, solid, prec)
    assert s.steps_per_launch()["in_use"]
    o = oracle_mod.Oracle(nx, ny, 0.9, 1e-5, -2e-6, solid, prec, "v0")
    for x in (s, o):
        x.set_populations(S)
    done = 0
    for n in (2, 3, 4):
        s.run(n)
        o.run(n)
        done += n
        assert_bitwise(s.populations(), o.populations(), f"{grid} after {done} steps")
    s.close()


@pytest.mark.parametrize("grid", [(1, 1), (1, 7), (5, 1), (3, 3), (64, 2), (257, 20), (130, 5)])
def test_fp16_two_step_degenerate_grids_equal_one_step(lbm, grid):
    """FP16 storage on the degenerate grids of the test above (a last 256-column tile one cell
    wide: 257 = 256 + 1): two steps per launch give the raw binary16 state of one step per launch
    bit for bit, after even and odd step counts (the odd start takes one single step first)."""
    nx, ny = grid
    solid = np.zeros(nx * ny, np.uint8)
    if nx >= 3 and ny >= 3:
        solid[(ny // 2) * nx + nx // 2] = 1
    rng = np.random.default_rng(nx * 137 + ny)
    S = (rng.standard_normal(nx * ny * 9) * 1e-3).astype(np.float32)
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    out = []
    for spl in (2, 1):
        s = lbm.Solver(nx, ny, 0.9, 1e-5, -2e-6, solid, "fp16")
        s.set_steps_per_launch(spl)
        assert s.steps_per_launch()["in_use"] == (spl == 2)
        s.set_populations(S)
        states = []
        for n in (2, 3, 4):
            s.run(n)
            states.append(s.get_state())
        out.append(states)
        s.close()
    for k, (a, b) in enumerate(zip(*out)):
        assert_bitwise(a, b, f"FP16 {grid} after run {k}")
