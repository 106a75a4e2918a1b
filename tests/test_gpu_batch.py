"""Batched persistent step: 256 independent domains of uneven sizes and parameters in one launch
per step, 100-step HIP graphs. Bar: every domain bit-identical to the single-domain step (and,
for small domains, to the oracle)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def single(lbm, d, steps, prec="float", init=(1.0, 0.0, 0.0)):
    s = lbm.Solver(d["nx"], d["ny"], d["tau"], d["fx"], d["fy"], d["solid"], prec)
    s.init_equilibrium(*init)
    s.run(steps)
    out = s.populations()
    s.close()
    return out


def test_256_domains_100_step_graph_bitwise(lbm):
    doms = lbm.random_domains(seed=0, n=256)
    b = lbm.Batch(doms, "float")
    b.set_graph_steps(100)
    b.run(100)
    assert b.total_cells() == sum(d["nx"] * d["ny"] for d in doms)
    bad = []
    for k, d in enumerate(doms):
        if np.count_nonzero(bits(b.populations(k)) != bits(single(lbm, d, 100))):
            bad.append(k)
    assert not bad, f"domains differing from the single-domain step: {bad[:10]}"


@pytest.mark.parametrize("prec", ("float", "double"))
def test_odd_graph_length_shapes_and_oracle(lbm, oracle_mod, prec):
    doms = lbm.random_domains(seed=5, n=12, nx_range=(20, 150), ny_range=(8, 70),
                              radius_range=(2, 8))
    init = (1.0, 0.01, -0.005)
    b = lbm.Batch(doms, prec, init=init)
    b.set_graph_steps(7)
    b.run(11)
    b.set_shape(64, 16, 5)
    b.run(6)
    b.set_shape(512, 1, 3)
    b.set_graph_steps(4)
    b.run(6)  # 23 steps in total: graphs of both parities and direct launches
    for k, d in enumerate(doms):
        ref = single(lbm, d, 23, prec, init)
        assert np.count_nonzero(bits(b.populations(k)) != bits(ref)) == 0, k
        if k < 4:
            o = oracle_mod.Oracle(d["nx"], d["ny"], d["tau"], d["fx"], d["fy"], d["solid"], prec, "v0")
            o.init_equilibrium(*init)
            o.run(23)
            assert np.count_nonzero(bits(o.populations()) != bits(ref)) == 0, k


@pytest.mark.parametrize("prec", ("float", "double"))
@pytest.mark.parametrize("order", (0, 1, 2))
def test_every_order_is_bit_identical(lbm, prec, order):
    """The three ways of walking the work (grid-stride tiles, XCD-banded chunks, wave work-list
    of 64-column segments) with several shapes: each domain equals its single-domain run."""
    doms = lbm.random_domains(seed=13, n=20, nx_range=(33, 400), ny_range=(17, 150),
                              radius_range=(2, 10))
    b = lbm.Batch(doms, prec)
    b.set_order(order, order != 1)
    assert b.order() == (order, order != 1)
    b.set_graph_steps(5)
    for shape, steps in (((256, 2, 64), 7), ((512, 4, 9), 6), ((128, 1, 300), 5)):
        b.set_shape(*shape)
        b.run(steps)
    for k, d in enumerate(doms):
        ref = single(lbm, d, 18, prec)
        assert np.count_nonzero(bits(b.populations(k)) != bits(ref)) == 0, (order, k)


def test_autotune_keeps_results(lbm):
    doms = lbm.random_domains(seed=9, n=16, nx_range=(32, 300), ny_range=(32, 200))
    b = lbm.Batch(doms, "float")
    tx, tr, wg, mlups = b.autotune(steps=2)
    assert tx % 64 == 0 and tr >= 1 and wg >= 1 and mlups > 0
    b.run(30)
    for k in (0, 7, 15):
        assert np.count_nonzero(bits(b.populations(k)) != bits(single(lbm, doms[k], 30))) == 0


def test_batch_rejects_invalid_domains(lbm):
    d = lbm.random_domains(seed=1, n=1, nx_range=(40, 40), ny_range=(40, 40))[0]
    bad = dict(d, tau=0.5)
    with pytest.raises(ValueError, match="tau"):
        lbm.Batch([d, bad])


@pytest.mark.parametrize("prec", ("float", "double"))
def test_two_step_batch_matches_one_step_batch(lbm, oracle_mod, prec):
    """Two steps per launch over the batch (per-domain sub-tiles, one launch per pair of steps,
    several sub-tiles of different domains and heights packed into one workgroup): sub-tile
    widths from one wave to wider than most domains (domains up to 1100 columns in two or
    three column tiles), bands of uneven heights, one workgroup per item and persistent grids,
    graphs of both parities, direct pairs and an odd remainder; then the modelled band height.
    Every domain equals the one-step batch and, for two, the oracle."""
    doms = lbm.random_domains(seed=21, n=30, nx_range=(20, 1100), ny_range=(9, 300),
                              radius_range=(2, 12))
    init = (1.0, 0.02, -0.01)
    out = []
    for spl in (2, 1):
        b = lbm.Batch(doms, prec, init=init)
        b.set_graph_steps(6)
        if spl == 2:
            # one workgroup per tile, and persistent grids smaller than the tile count
            for tile, wg, steps in (((64, 16), 0, 4), ((192, 48), 7, 5), ((512, 64), 300, 7),
                                    ((320, 100), 0, 2), ((128, 5), 0, 2), ((448, -1), 0, 2)):
                if prec == "double":  # its LDS ring holds twice the bytes: at most 256 columns
                    tile = (min(tile[0], 256), tile[1])
                b.set_steps_per_launch(2, *tile, wg)
                lay = b.two_step_layout()
                rows = lay["band_rows"]
                assert lay["modelled"] == (tile[1] == -1)
                assert rows == tile[1] if tile[1] > 0 else 16 <= rows <= 96
                assert b.steps_per_launch() == {"steps_per_launch": 2, "tile_x": tile[0],
                                                "tile_rows": rows, "workgroups": wg}
                assert lay["threads"] == (256 if prec == "double" else 512)
                assert 0.3 < lay["fill"] <= 1.0 and lay["items"] >= 1
                b.run(steps)
        else:
            b.set_steps_per_launch(1)
            assert b.steps_per_launch()["steps_per_launch"] == 1
            b.run(22)
        out.append([b.populations(k) for k in range(len(doms))])
        b.close()
    for k in range(len(doms)):
        assert np.count_nonzero(bits(out[0][k]) != bits(out[1][k])) == 0, k
    for k in (0, 1):
        d = doms[k]
        o = oracle_mod.Oracle(d["nx"], d["ny"], d["tau"], d["fx"], d["fy"], d["solid"], prec, "v0")
        o.init_equilibrium(*init)
        o.run(22)
        assert np.count_nonzero(bits(o.populations()) != bits(out[0][k])) == 0, k
