"""Captured plans (pimdal_plan_*) and the asynchronous Q1 / Q6 forms, on the
GPU against the CPU restatement and the eager calls.  Every graph is
replayed several times, with eager work and synchronisations between the
replays, and every replay is checked."""
import ctypes as C

import numpy as np
import pytest

from pimdal_amd import _lib, datagen, operators as ops
from tests.test_gpu_aggregate import _cmp_q1

pytestmark = pytest.mark.gpu
T = datagen.SELECT_THRESHOLD


def _eager_noise(ctx, orc):
    """Eager work between replays: a sort of other data, checked, then a sync."""
    k, v = orc.sort_workload(1 << 18, seed=77)
    ok, ov = ops.sort_pairs(ctx, ctx.to_device(k), ctx.to_device(v))
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(ok.numpy(), ek) and np.array_equal(ov.numpy(), ev)
    ctx.sync()


def test_plan_select_then_sort_replays(ctx, orc):
    """select_dev and sort_pairs_dev (whole key) captured as one plan; between
    replays the inputs are rewritten in place, and each replay's outputs are
    those of the data present at its launch."""
    n = 1 << 20
    dk, dp = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32)
    op_, orow, cnt = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32), ctx.empty(1, np.uint64)
    sk, sv = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32)
    with ops.Plan(ctx) as plan:
        _lib.call("pimdal_select_dev", ctx.handle, _lib.PIMDAL_U32, dk.vp, dp.vp, n, 0, T, op_.vp, orow.vp, cnt.vp)
        _lib.call("pimdal_sort_pairs_dev", ctx.handle, dk.vp, dp.vp, 4, n, sk.vp, sv.vp, 0, 32)
    try:
        for rep in range(3):
            k, p = orc.select_workload(n, seed=300 + rep)
            _lib.call("pimdal_memcpy_h2d", ctx.handle, dk.vp, k.ctypes.data_as(C.c_void_p), k.nbytes)
            _lib.call("pimdal_memcpy_h2d", ctx.handle, dp.vp, p.ctypes.data_as(C.c_void_p), p.nbytes)
            plan.launch()
            ctx.sync()
            c = int(cnt.numpy()[0])
            ep, er = orc.select(k, p, 0, T)
            assert c == len(ep) and np.array_equal(op_.numpy(c), ep) and np.array_equal(orow.numpy(c), er)
            ek, ev = orc.sort_pairs(k, p)
            assert np.array_equal(sk.numpy(), ek) and np.array_equal(sv.numpy(), ev)
            _eager_noise(ctx, orc)
    finally:
        plan.destroy()


@pytest.mark.parametrize("by_key", [False, True], ids=["flags", "orderkey"])
@pytest.mark.parametrize("n", [1, 100_003, 1 << 20])
def test_q1_async_eager_matches_oracle(ctx, orc, by_key, n):
    t = orc.fact_table(n)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    out, ng, st = ops.filter_groupby_q1_async(ctx, d, 2505, by_key=by_key)
    ctx.sync()
    status, got = ops.q1_to_host(out, ng, st)
    assert status == 0
    exp = orc.q1_key(t, 2505) if by_key else orc.q1_flags(t, 2505)
    got.pop("groups")
    _cmp_q1(got, exp, ordered=not by_key)


@pytest.mark.parametrize("by_key", [False, True], ids=["flags", "orderkey"])
def test_q1_plan_replays_6e7_rows(ctx, orc, by_key):
    """The filter + group-by pipeline of 6e7 rows as one captured plan (flags:
    domain, codes, dense pass, reduction; by key: pack, hash partition passes,
    per-partition LDS tables), replayed three times with eager work between;
    every replay equals the oracle and the eager synchronous call."""
    n = 60_000_000
    d = datagen.fact_table(ctx, n)
    t = {c: v.numpy() for c, v in d.items()}
    exp = orc.q1_key(t, 2505) if by_key else orc.q1_flags(t, 2505)
    eager = ops.filter_groupby_q1(ctx, d, 2505, by_key=by_key)
    _cmp_q1(eager, exp, ordered=not by_key)
    with ops.Plan(ctx) as plan:
        out, ng, st = ops.filter_groupby_q1_async(ctx, d, 2505, by_key=by_key)
    try:
        for _ in range(3):
            for f in ops.Q1_FIELDS:   # stale results must not survive a replay
                _lib.call("pimdal_dev_copy", ctx.handle, out[f].vp, d["quantity"].vp, 8)
            plan.launch()
            ctx.sync()
            status, got = ops.q1_to_host(out, ng, st)
            assert status == 0 and got.pop("groups") == len(exp["key"])
            _cmp_q1(got, exp, ordered=not by_key)
            _eager_noise(ctx, orc)
    finally:
        plan.destroy()


def test_q1_async_flags_beyond_dense_path_reports_unsupported(ctx, orc):
    n = 200_000
    t = orc.fact_table(n)
    t["flag1"] = orc.gen_range(np.uint8, 3, 17, n, 60, 90)   # 31 x 2 combinations
    d = {c: ctx.to_device(v) for c, v in t.items()}
    out, ng, st = ops.filter_groupby_q1_async(ctx, d, 2000)
    ctx.sync()
    assert int(st.numpy()[0]) == -2 and int(ng.numpy()[0]) == 0


@pytest.mark.parametrize("col,val", [("price", 1 << 31), ("quantity", 32768)])
def test_q1_async_key_beyond_packed_width_reports_unsupported(ctx, orc, col, val):
    n = 120_000
    t = orc.fact_table(n, groups=40_000)
    t["date"][777] = 0
    t[col][777] = val
    d = {c: ctx.to_device(v) for c, v in t.items()}
    out, ng, st = ops.filter_groupby_q1_async(ctx, d, 2505, by_key=True)
    ctx.sync()
    assert int(st.numpy()[0]) == -2


@pytest.mark.parametrize("by_key", [False, True], ids=["flags", "orderkey"])
def test_q1_async_capacity_short(ctx, orc, by_key):
    """More groups than capacity: PIMDAL_E_CAPACITY, the exact count, and
    nothing written past the caller's arrays (a guard region after them)."""
    n = 100_000
    t = orc.fact_table(n, groups=5_000)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    exp = orc.q1_key(t, 2505) if by_key else orc.q1_flags(t, 2505)
    cap = 3 if not by_key else 1000
    guard = 64
    out = {f: ctx.empty((cap + guard) * (2 if f == "sum_charge" else 1), np.int32 if f == "key" else np.int64)
           for f in ops.Q1_FIELDS}
    sentinel = {f: np.full(out[f].n, -7, dtype=out[f].dtype) for f in ops.Q1_FIELDS}
    for f in ops.Q1_FIELDS:
        _lib.call("pimdal_memcpy_h2d", ctx.handle, out[f].vp, sentinel[f].ctypes.data_as(C.c_void_p),
                  sentinel[f].nbytes)
    res = ops._Q1Result(*[out[f].ptr for f in ops.Q1_FIELDS])
    ng, st = ctx.empty(1, np.uint64), ctx.empty(1, np.int32)
    _lib.call("pimdal_filter_groupby_q1_dev_async", ctx.handle, d["date"].vp, d["flag1"].vp, d["flag2"].vp,
              d["orderkey"].vp if by_key else None, d["quantity"].vp, d["price"].vp, d["discount"].vp, d["tax"].vp,
              n, 2505, C.byref(res), cap, ng.vp, st.vp)
    ctx.sync()
    assert int(st.numpy()[0]) == -5 and int(ng.numpy()[0]) == len(exp["key"])
    for f in ops.Q1_FIELDS:
        w = 2 if f == "sum_charge" else 1
        tail = out[f].numpy()[cap * w:]
        assert np.all(tail == -7), f
    keys = out["key"].numpy(cap)
    assert set(keys.tolist()) <= set(np.asarray(exp["key"]).tolist())


def test_q1_sync_capacity_short_writes_within_capacity(ctx, orc):
    """The synchronous flag form with capacity below the group count returns
    PIMDAL_E_CAPACITY without writing past the caller's arrays."""
    n = 50_000
    t = orc.fact_table(n)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    cap, guard = 2, 32
    out = {f: ctx.empty((cap + guard) * (2 if f == "sum_charge" else 1), np.int32 if f == "key" else np.int64)
           for f in ops.Q1_FIELDS}
    for f in ops.Q1_FIELDS:
        s = np.full(out[f].n, -7, dtype=out[f].dtype)
        _lib.call("pimdal_memcpy_h2d", ctx.handle, out[f].vp, s.ctypes.data_as(C.c_void_p), s.nbytes)
    res = ops._Q1Result(*[out[f].ptr for f in ops.Q1_FIELDS])
    g = C.c_uint64()
    rc = _lib.load().pimdal_filter_groupby_q1_dev(ctx.handle, d["date"].vp, d["flag1"].vp, d["flag2"].vp, None,
                                                  d["quantity"].vp, d["price"].vp, d["discount"].vp, d["tax"].vp, n,
                                                  2505, C.byref(res), cap, C.byref(g))
    assert rc == -5
    for f in ops.Q1_FIELDS:
        w = 2 if f == "sum_charge" else 1
        assert np.all(out[f].numpy()[cap * w:] == -7), f


def test_q6_async_in_a_plan(ctx, orc):
    n = 3_000_001
    t = orc.fact_table(n)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    with ops.Plan(ctx) as plan:
        rev, rows = ops.q6_async(ctx, d)
    try:
        for _ in range(3):
            plan.launch()
            ctx.sync()
            assert (int(rev.numpy()[0]), int(rows.numpy()[0])) == orc.q6(t)
            _eager_noise(ctx, orc)
    finally:
        plan.destroy()


def test_plan_refuses_a_synchronous_call(ctx, orc):
    """A call that reads results back on the host fails inside a capture, and
    the plan then has nothing to launch; the context works as before."""
    rk, rp, sk, sp = orc.join_workload(1 << 12, 1 << 14)
    dv = [ctx.to_device(a) for a in (rk, rp, sk, sp)]
    op_, ob = ctx.empty(1 << 14, np.uint32), ctx.empty(1 << 14, np.uint32)
    h = C.c_void_p()
    _lib.call("pimdal_plan_begin", ctx.handle, C.byref(h))
    cnt = C.c_uint64()
    rc = _lib.load().pimdal_hash_join_dev(ctx.handle, dv[0].vp, dv[1].vp, rk.size, dv[2].vp, dv[3].vp, sk.size,
                                          op_.vp, ob.vp, 1 << 14, C.byref(cnt))
    assert rc == -2 and "plan" in _lib.last_error()
    assert _lib.load().pimdal_plan_end(h) != 0
    assert _lib.load().pimdal_plan_launch(h) != 0
    _lib.call("pimdal_plan_destroy", h)
    gp, gb, k = ops.hash_join(ctx, *dv)
    ep, eb = orc.hash_join(rk, rp, sk, sp)
    assert k == len(ep)


def test_plan_sort_detecting_bits_is_refused(ctx, orc):
    """Detecting the varying key bits reads a word back: not capturable."""
    n = 1 << 16
    dk, dv = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32)
    ok, ov = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32)
    h = C.c_void_p()
    _lib.call("pimdal_plan_begin", ctx.handle, C.byref(h))
    rc = _lib.load().pimdal_sort_pairs_dev(ctx.handle, dk.vp, dv.vp, 4, n, ok.vp, ov.vp, -1, -1)
    assert rc == -2
    assert _lib.load().pimdal_plan_end(h) != 0
    _lib.call("pimdal_plan_destroy", h)
