"""lbm_macroscopic_noalloc (include/lbm_hip.h): lbm_macroscopic without device allocation, the
fields staged in the library's fixed 12 MiB device array. Called through the C ABI; its outputs
must equal lbm_macroscopic's bit for bit on two calls in a row, the second on a smaller grid,
and a grid whose three fields exceed the array must be rejected before any device work."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import walls_block_mask

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 12 << 20


def call(fn, s):
    out = tuple(np.empty(s.N, dtype=s.dtype) for _ in range(3))
    rc = fn(s.handle, *(a.ctypes.data for a in out))
    return rc, out


def same_bits(a, b):
    ia = a.view(np.uint64 if a.dtype == np.float64 else np.uint32)
    ib = b.view(np.uint64 if b.dtype == np.float64 else np.uint32)
    return int(np.count_nonzero(ia != ib))


@pytest.mark.parametrize("prec", ("float", "double", "fp16"))
def test_noalloc_fields_equal_macroscopic_on_two_calls(lbm, prec):
    print(f"no-alloc fields, {prec}: 1024x256 cylinder, then 53x31 walls + block", flush=True)
    L = lbm.lib()
    nx, ny = 1024, 256
    solid = np.maximum(lbm.make_channel(nx, ny), lbm.make_cylinder(nx, ny, 256.0, 129.0, 32.0))
    big = lbm.Solver(nx, ny, 0.6, 1e-5, 0.0, solid, prec)
    big.init_equilibrium(1.0, 0.03, 0.0)
    big.run(30)
    small = lbm.Solver(53, 31, 0.8, 1e-5, 2e-6, walls_block_mask(), prec)
    small.init_equilibrium(1.0, 0.02, 0.01)
    small.run(21)
    # the entry point twice in a row: the second call's grid is smaller than the first's, so
    # the array still holds the first call's values past the second call's fields
    rc1, got_big = call(L.lbm_macroscopic_noalloc, big)
    assert rc1 == lbm.LBM_OK, lbm.last_error()
    rc2, got_small = call(L.lbm_macroscopic_noalloc, small)
    assert rc2 == lbm.LBM_OK, lbm.last_error()
    for s, got in ((big, got_big), (small, got_small)):
        rc, ref = call(L.lbm_macroscopic, s)
        assert rc == lbm.LBM_OK, lbm.last_error()
        for a, b in zip(got, ref):
            assert same_bits(a, b) == 0
    # fields of a non-trivial flow, not a zero state
    assert np.abs(got_big[1]).max() > 0.01
    assert np.abs(got_small[2]).max() > 0.0
    # after more steps the array is reused with the new state
    small.run(5)
    rc, again = call(L.lbm_macroscopic_noalloc, small)
    assert rc == lbm.LBM_OK, lbm.last_error()
    _, ref = call(L.lbm_macroscopic, small)
    for a, b in zip(again, ref):
        assert same_bits(a, b) == 0
    big.close()
    small.close()


@pytest.mark.parametrize("nx,ny,prec,ok", [
    (1024, 1024, "float", True),    # exactly 12 MiB of fields
    (1024, 1024, "double", False),  # 24 MiB
    (2048, 1024, "float", False),
    (1024, 512, "double", True),    # exactly 12 MiB
])
def test_noalloc_limit(lbm, nx, ny, prec, ok):
    print(f"no-alloc limit: {nx}x{ny} {prec}", flush=True)
    L = lbm.lib()
    s = lbm.Solver(nx, ny, 0.8, 1e-6, 0.0, lbm.make_channel(nx, ny), prec)
    s.init_equilibrium(1.0, 0.01, 0.0)
    s.run(2)
    rc, got = call(L.lbm_macroscopic_noalloc, s)
    if ok:
        assert rc == lbm.LBM_OK, lbm.last_error()
        _, ref = call(L.lbm_macroscopic, s)
        for a, b in zip(got, ref):
            assert same_bits(a, b) == 0
    else:
        assert rc == lbm.LBM_EINVAL
        assert "maximum is 12582912" in lbm.last_error()
        with pytest.raises(ValueError, match="scratch"):
            s.macroscopic_noalloc()
    s.close()
