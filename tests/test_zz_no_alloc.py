"""pimdal_select_dev_noalloc (status words in the library's static device
array) against pimdal_select_dev (status words from the pool), called
through the C ABI: identical payloads, row ids and counts on two calls in a
row, the second smaller, and no pool memory taken by the no-allocation
form."""
import ctypes as C

import numpy as np
import pytest

from pimdal_amd import _lib, datagen

pytestmark = pytest.mark.gpu
T = datagen.SELECT_THRESHOLD


def _select(ctx, fn, dk, dp, n, lo, hi, out_p, out_r, cnt):
    _lib.call(fn, ctx.handle, _lib.PIMDAL_U32, dk.vp, dp.vp, n, lo, hi, out_p.vp, out_r.vp, cnt.vp)
    ctx.sync()
    k = int(cnt.numpy()[0])
    return k, out_p.numpy(k), out_r.numpy(k)


@pytest.mark.parametrize("block", [1024, 256])
def test_noalloc_select_matches_pool_select(ctx, orc, block):
    ctx.set_tune("select_block", block)
    try:
        for n, seed in ((3_000_017, 91), (100_003, 92)):   # second call smaller than the first
            k, p = orc.select_workload(n, seed=seed)
            dk, dp = ctx.to_device(k), ctx.to_device(p)
            ref = [ctx.empty(n, np.uint32), ctx.empty(n, np.uint32), ctx.empty(1, np.uint64)]
            got = [ctx.empty(n, np.uint32), ctx.empty(n, np.uint32), ctx.empty(1, np.uint64)]
            rk, rp, rr = _select(ctx, "pimdal_select_dev", dk, dp, n, 0, T, *ref)
            before = ctx.mem_stats()
            gk, gp, gr = _select(ctx, "pimdal_select_dev_noalloc", dk, dp, n, 0, T, *got)
            after = ctx.mem_stats()
            assert gk == rk and np.array_equal(gp, rp) and np.array_equal(gr, rr)
            assert (after["live"], after["reserved"]) == (before["live"], before["reserved"])
            ep, er = orc.select(k, p, 0, T)
            assert np.array_equal(gp, ep) and np.array_equal(gr, er)
    finally:
        ctx.set_tune("select_block", 1024)


def test_noalloc_select_empty_and_unaligned(ctx, orc):
    k, p = orc.select_workload(50_001, seed=7)
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    out = [ctx.empty(50_001, np.uint32), ctx.empty(50_001, np.uint32), ctx.empty(1, np.uint64)]
    kk, _, _ = _select(ctx, "pimdal_select_dev_noalloc", dk, dp, 50_001, 5, 5, *out)
    assert kk == 0
    vk, vp = dk.view(1, 49_999), dp.view(2, 49_999)   # 4-byte aligned only: the 256-thread kernel
    gk, gp, gr = _select(ctx, "pimdal_select_dev_noalloc", vk, vp, 49_999, 0, T, *out)
    ep, er = orc.select(k[1:50_000], p[2:50_001], 0, T)
    assert gk == len(ep) and np.array_equal(gp, ep) and np.array_equal(gr, er)


def test_noalloc_select_rejects_oversized_input(ctx):
    d = ctx.empty(16, np.uint32)
    c = ctx.empty(1, np.uint64)
    rc = _lib.load().pimdal_select_dev_noalloc(ctx.handle, _lib.PIMDAL_U32, d.vp, d.vp, (1 << 29) + 1, 0, T,
                                                d.vp, d.vp, c.vp)
    assert rc == -5 and "2^29" in _lib.last_error()
