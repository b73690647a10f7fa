"""Hash / sort aggregation and the fused Q1 group-by on the GPU vs the CPU restatement."""
import numpy as np
import pytest

import pimdal_amd
from pimdal_amd import operators as ops, datagen
from tests.util import canon

pytestmark = pytest.mark.gpu


def agg(ctx, fn, k, v):
    ok, os_, oc, g = fn(ctx, ctx.to_device(k), ctx.to_device(v))
    return canon(ok.numpy(g), os_.numpy(g), oc.numpy(g))


@pytest.mark.parametrize("which", ["hash", "sort"])
@pytest.mark.parametrize("n,groups", [(1, 1), (1000, 50), (500_000, 50), (300_001, 20_000), (200_000, 200_000),
                                      (400_000, 1 << 32)])
def test_aggr_matches_oracle(ctx, orc, which, n, groups):
    k = orc.gen_range(np.uint32, n, 1, n, 0, groups - 1)
    v = orc.gen_range(np.int64, n, 2, n, -(1 << 40), 1 << 40)
    fn = ops.hash_aggr if which == "hash" else ops.sort_aggr
    got = agg(ctx, fn, k, v)
    exp = orc.hash_aggr(k, v)
    assert all(np.array_equal(a, b) for a, b in zip(got, exp))


def test_aggr_reference_shape(ctx, orc):
    # reference/db_operators/aggregate.cpp:33-36: keys on [1, 50], values on [1, 10]
    k = orc.gen_range(np.uint32, 1, 1, 2_000_000, 1, 50)
    v = orc.gen_range(np.int64, 1, 2, 2_000_000, 1, 10)
    for fn in (ops.hash_aggr, ops.sort_aggr):
        got = agg(ctx, fn, k, v)
        exp = orc.hash_aggr(k, v)
        assert all(np.array_equal(a, b) for a, b in zip(got, exp))


def test_sort_aggr_zipf_skew(ctx, orc):
    k, _ = orc.sort_workload(1 << 20, zipf=True, distinct=1000)
    v = orc.gen_range(np.int64, 5, 2, 1 << 20, 0, 1000)
    for fn in (ops.sort_aggr, ops.hash_aggr):
        got = agg(ctx, fn, k, v)
        exp = orc.hash_aggr(k, v)
        assert all(np.array_equal(a, b) for a, b in zip(got, exp))


AGGR_SHAPES = {"k50": (1, 50), "k4000": (1, 4000), "k25": (1, 25), "k2^24": (0, (1 << 24) - 1)}


@pytest.mark.parametrize("shape", sorted(AGGR_SHAPES))
def test_aggr_bench_workloads_match_oracle(ctx, orc, shape):
    """The benchmark's group-by inputs, generated on the device, equal the
    host recipe; hash (canonical order) and sort (its own order) outputs
    equal the oracle's at 2^22 rows."""
    lo, hi = AGGR_SHAPES[shape]
    n = 1 << 22
    dk, dv = datagen.aggr_workload(ctx, n, lo, hi)
    k, v = orc.aggr_workload(n, lo, hi)
    assert np.array_equal(dk.numpy(), k) and np.array_equal(dv.numpy(), v)
    exp = orc.sort_aggr(k, v)
    for fn in (ops.hash_aggr, ops.sort_aggr):
        ok, os_, oc, g = fn(ctx, dk, dv)
        got = (ok.numpy(g), os_.numpy(g), oc.numpy(g))
        if fn is ops.hash_aggr:
            got = canon(*got)
        assert all(np.array_equal(a, b) for a, b in zip(got, exp)), fn.__name__


@pytest.mark.parametrize("shape", ["k50", "k2^24"])
def test_aggr_2p27_hash_and_sort_agree(ctx, shape):
    """At the benchmark's 2^27 rows (too slow for the oracle with 2^24 keys):
    hash and sort outputs hold the same groups, sort's in strictly ascending
    key order, counts summing to the row count and sums to the value total."""
    lo, hi = AGGR_SHAPES[shape]
    n = 1 << 27
    dk, dv = datagen.aggr_workload(ctx, n, lo, hi)
    total = int(dv.numpy().sum())
    hk, hs, hc, hg = ops.hash_aggr(ctx, dk, dv)
    sk, ss, sc, sg = ops.sort_aggr(ctx, dk, dv)
    assert hg == sg
    s = (sk.numpy(sg), ss.numpy(sg), sc.numpy(sg))
    assert np.all(np.diff(s[0].astype(np.int64)) > 0)
    assert int(s[2].sum()) == n and int(s[1].sum()) == total
    h = canon(hk.numpy(hg), hs.numpy(hg), hc.numpy(hg))
    assert all(np.array_equal(a, b) for a, b in zip(h, s))
    for a in (dk, dv, hk, hs, hc, sk, ss, sc):
        a.free()


def _both_paths(ctx, orc, k, v, small_fits=None):
    """hash and sort aggregation, with and without the few-groups LDS path,
    and with and without the dense key-range path, against the oracle (sort
    output in its own order, hash canonical)."""
    exp = orc.sort_aggr(k, v)
    dk, dv = ctx.to_device(k), ctx.to_device(v)
    try:
        for direct, small, dense in ((1, 1, 1), (0, 1, 1), (0, 0, 1), (0, 0, 0)):
            ctx.set_tune("aggr_direct", direct)
            ctx.set_tune("aggr_small", small)
            if dense:
                ctx.clear_tune("aggr_dense_max")
            else:
                ctx.set_tune("aggr_dense_max", 0)
            for fn in (ops.hash_aggr, ops.sort_aggr):
                ok, os_, oc, g = fn(ctx, dk, dv)
                got = (ok.numpy(g), os_.numpy(g), oc.numpy(g))
                if fn is ops.hash_aggr:
                    got = canon(*got)
                assert g == exp[0].size, (fn.__name__, direct, small, dense)
                assert all(np.array_equal(a, b) for a, b in zip(got, exp)), (fn.__name__, direct, small, dense)
    finally:
        ctx.clear_tune("aggr_direct")
        ctx.clear_tune("aggr_small")
        ctx.clear_tune("aggr_dense_max")


def test_aggr_paths_max_key_and_ragged(ctx, orc):
    """Key 0xFFFFFFFF (the LDS table's empty marker) and 0 among 1000 keys,
    row counts around the 8192-row round and tile."""
    for n in (1, 2, 8191, 8192, 8193, 1_000_003):
        k = orc.gen_range(np.uint32, 31, 1, n, 0, 999)
        k[::7] = 0xFFFFFFFF
        k[3::11] = 0
        v = orc.gen_range(np.int64, 31, 2, n, -(1 << 40), 1 << 40)
        _both_paths(ctx, orc, k, v)


def test_aggr_paths_block_table_limit(ctx, orc):
    """One block's rows (8192 rows: one block) with 4096 distinct keys fit its
    table; 4097 do not and the many-groups path takes over."""
    for distinct in (4096, 4097):
        k = (np.arange(8192, dtype=np.uint64) % distinct).astype(np.uint32) * np.uint32(2654435761)
        v = orc.gen_range(np.int64, 32, 2, 8192, 1, 10)
        _both_paths(ctx, orc, k, v)


def test_aggr_paths_hbm_table_limit(ctx, orc):
    """Sorted keys: every block meets a few hundred keys, all blocks together
    2^16 (past the HBM table's 32768): the many-groups path takes over; and
    32768 keys exactly, which fit."""
    for distinct in (1 << 16, 1 << 15):
        n = 1 << 21
        k = np.repeat(np.arange(distinct, dtype=np.uint32) * np.uint32(7919), n // distinct)
        v = orc.gen_range(np.int64, 33, 2, n, -5, 5)
        _both_paths(ctx, orc, k, v)


@pytest.mark.parametrize("items", [4, 8, 16])
def test_sort_aggr_runs_across_tiles(ctx, orc, items):
    """Sorted-run reduce (4, 8 or 16 rows per thread: tiles of 4096 to 16384
    rows) with runs spanning thousands of tiles (few keys, the LDS path off),
    Zipf runs, runs ending on tile edges, and a ragged last tile."""
    n = (1 << 22) + 1000
    ctx.set_tune("as_runs_items", items)
    try:
        for k in (orc.gen_range(np.uint32, 34, 1, n, 1, 25),
                  orc.sort_workload(n, zipf=True, distinct=1 << 16)[0],
                  np.repeat(np.arange(n // 4096 + 1, dtype=np.uint32), 4096)[:n]):
            v = orc.gen_range(np.int64, 34, 2, n, -(1 << 50), 1 << 50)
            _both_paths(ctx, orc, k, v)
    finally:
        ctx.clear_tune("as_runs_items")


def _dense_ran(ctx, fn, dk, dv):
    ctx.set_profiling(True)
    ctx.reset_stats()
    try:
        ok, os_, oc, g = fn(ctx, dk, dv)
        return (ok.numpy(g), os_.numpy(g), oc.numpy(g)), ctx.kernel_stat("ad_sum")[0] > 0
    finally:
        ctx.set_profiling(False)


@pytest.mark.parametrize("case", ["span_2^24", "span_2^24_plus_1", "top_of_u32", "from_2^31", "ragged_8193",
                                  "one_key_per_partition_edge", "outlier_below_sample", "wrap_low_outlier"])
def test_aggr_dense_key_range(ctx, orc, case):
    """The key-range path: ranges of exactly 2^24 keys (2048 partitions) and
    one more (the hashed path instead), keys ending at 0xFFFFFFFF, starting
    at 2^31, a ragged row count, and keys on partition edges.  The window
    comes from a 2^16-key sample (rows 0, 32, 64, ... here): a key the sample
    missed below its window sends the call to the exact min/max (still the
    key-range path), and so does a key near 0 under a window near the top
    of the key space whose partitions run past 2^32 (then the hashed path:
    the span is too wide)."""
    n = 1 << 21
    v = orc.gen_range(np.int64, 36, 2, n, -(1 << 45), 1 << 45)
    if case == "span_2^24":
        k = orc.gen_range(np.uint32, 36, 1, n, 5, 5 + (1 << 24) - 1)
        k[0], k[1] = 5, 5 + (1 << 24) - 1
    elif case == "span_2^24_plus_1":
        k = orc.gen_range(np.uint32, 36, 1, n, 5, 5 + (1 << 24))
        k[0], k[1] = 5, 5 + (1 << 24)
    elif case == "top_of_u32":
        k = orc.gen_range(np.uint32, 36, 1, n, 0xFFFFFFFF - (1 << 20), 0xFFFFFFFF)
        k[::97] = 0xFFFFFFFF
    elif case == "from_2^31":
        k = orc.gen_range(np.uint32, 36, 1, n, 1 << 31, (1 << 31) + 300_000)
    elif case == "ragged_8193":
        n = 8193
        k = orc.gen_range(np.uint32, 36, 1, n, 0, 1 << 20)
        v = v[:n]
    elif case == "outlier_below_sample":
        k = orc.gen_range(np.uint32, 36, 1, n, 1 << 20, 1 << 21)
        k[1] = 100
    elif case == "wrap_low_outlier":
        k = orc.gen_range(np.uint32, 36, 1, n, 0xFFFFFFFF - (1 << 20), 0xFFFFFFFF)
        k[1] = 3
    else:
        base = np.arange(n, dtype=np.uint64) % 2048
        k = (base * np.uint64(8192) + np.uint64(8191) * (np.arange(n) % 2).astype(np.uint64)).astype(np.uint32)
    exp = orc.sort_aggr(k, v)
    dk, dv = ctx.to_device(k), ctx.to_device(v)
    ctx.set_tune("aggr_small", 0)   # straight to the many-groups paths
    try:
        for fn in (ops.hash_aggr, ops.sort_aggr):
            got, dense = _dense_ran(ctx, fn, dk, dv)
            assert dense == (case not in ("span_2^24_plus_1", "wrap_low_outlier")), (case, fn.__name__)
            if fn is ops.hash_aggr:
                got = canon(*got)
            assert all(np.array_equal(a, b) for a, b in zip(got, exp)), (case, fn.__name__)
    finally:
        ctx.clear_tune("aggr_small")


def test_aggr_capacity_short(ctx, orc):
    from pimdal_amd import _lib
    k = orc.gen_range(np.uint32, 35, 1, 100_000, 0, 99)
    v = orc.gen_range(np.int64, 35, 2, 100_000, 1, 10)
    dk, dv = ctx.to_device(k), ctx.to_device(v)
    try:
        for small in (1, 0):
            ctx.set_tune("aggr_small", small)
            for fn in (ops.hash_aggr, ops.sort_aggr):
                with pytest.raises(_lib.PimdalError, match="capacity"):
                    fn(ctx, dk, dv, capacity=99)
                ok, os_, oc, g = fn(ctx, dk, dv, capacity=100)
                assert g == 100
    finally:
        ctx.clear_tune("aggr_small")


def _cmp_q1(got, exp, ordered):
    keys = list(got["key"])
    if not ordered:
        order = np.argsort(got["key"], kind="stable")
        got = {f: (np.asarray(v)[order] if f != "sum_charge" else [v[i] for i in order]) for f, v in got.items()}
    assert list(got["key"]) == list(exp["key"])
    for f in ("sum_qty", "sum_price", "sum_disc", "sum_disc_price", "count"):
        assert np.array_equal(np.asarray(got[f]), np.asarray(exp[f])), f
    assert list(got["sum_charge"]) == list(exp["sum_charge"])


@pytest.mark.parametrize("n", [1, 1000, 1 << 20])
def test_q1_flags_dense(ctx, orc, n):
    t = orc.fact_table(n)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505)
    exp = orc.q1_flags(t, 2505)
    _cmp_q1(got, exp, ordered=True)


def test_q1_flags_many_groups_general_path(ctx, orc):
    n = 200_000
    t = orc.fact_table(n)
    t["flag1"] = orc.gen_range(np.uint8, 3, 17, n, 60, 90)   # 31 x 2 groups: beyond the dense path
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2000, capacity=4096)
    exp = orc.q1_flags(t, 2000)
    _cmp_q1(got, exp, ordered=False)


def test_q1_int128_charge(ctx, orc):
    n = 4096
    t = {"date": np.zeros(n, np.int32), "flag1": np.zeros(n, np.uint8), "flag2": np.ones(n, np.uint8),
         "quantity": np.ones(n, np.int64), "price": np.full(n, 90_000_000_000_000_000, np.int64),
         "discount": np.zeros(n, np.int64), "tax": np.full(n, 8, np.int64),
         "orderkey": np.full(n, 5, np.int32)}
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 0)
    assert got["sum_charge"][0] == n * 90_000_000_000_000_000 * 100 * 108
    got = ops.filter_groupby_q1(ctx, d, 0, by_key=True)
    assert got["sum_charge"][0] == n * 90_000_000_000_000_000 * 100 * 108


@pytest.mark.parametrize("n,groups", [(100_000, 25_000), (1 << 21, 1 << 19)])
def test_q1_by_orderkey(ctx, orc, n, groups):
    t = orc.fact_table(n, groups=groups)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505, by_key=True)
    exp = orc.q1_key(t, 2505)
    _cmp_q1(got, exp, ordered=False)


def test_q1_6e7_rows_full_size(ctx, orc):
    """Q1 at full size (6e7 rows): both variants vs the CPU restatement."""
    n = 60_000_000
    d = datagen.fact_table(ctx, n)
    t = {c: v.numpy() for c, v in d.items()}
    _cmp_q1(ops.filter_groupby_q1(ctx, d, 2505), orc.q1_flags(t, 2505), ordered=True)
    _cmp_q1(ops.filter_groupby_q1(ctx, d, 2505, by_key=True), orc.q1_key(t, 2505), ordered=False)


def test_q1_by_orderkey_signed_packed_fields(ctx, orc):
    """Negative measures that fit the packed widths (int16 quantity, int32
    price, int8 discount and tax) must sign-extend back exactly."""
    n = 150_000
    t = orc.fact_table(n, groups=30_000)
    t["quantity"] = orc.gen_range(np.int64, 21, 12, n, -32768, 32767)
    t["price"] = orc.gen_range(np.int64, 21, 13, n, -(1 << 31), (1 << 31) - 1)
    t["discount"] = orc.gen_range(np.int64, 21, 14, n, -128, 127)
    t["tax"] = orc.gen_range(np.int64, 21, 15, n, -128, 127)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505, by_key=True)
    _cmp_q1(got, orc.q1_key(t, 2505), ordered=False)


@pytest.mark.parametrize("col,val", [("quantity", 32768), ("quantity", -32768), ("price", 1 << 31), ("discount", -129),
                                     ("tax", 128)])
def test_q1_by_orderkey_one_row_beyond_packed_width(ctx, orc, col, val):
    """One kept row outside the packed widths (or with the reserved quantity
    -32768) sends the whole call down the unpacked path; the result must not
    change."""
    n = 120_000
    t = orc.fact_table(n, groups=40_000)
    t["date"][777] = 0
    t[col][777] = val
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505, by_key=True)
    _cmp_q1(got, orc.q1_key(t, 2505), ordered=False)


@pytest.mark.parametrize("sign", [1, -1])
def test_q1_by_orderkey_extreme_packed_charge_sums(ctx, orc, sign):
    """Groups of about 60000 rows at the extremes of the packed widths: the
    128-bit charge sums (about 6.7e15, negative for sign -1) of LDS tables,
    carries and sign words included."""
    n = 120_000
    t = orc.fact_table(n, groups=2)
    t["price"][:] = (1 << 31) - 1 if sign > 0 else -(1 << 31)
    t["discount"][:] = -128
    t["tax"][:] = 127
    t["quantity"][:] = 32767
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505, by_key=True)
    _cmp_q1(got, orc.q1_key(t, 2505), ordered=False)


def test_q1_by_orderkey_charge_beyond_int64_in_one_group(orc):
    """One group of 100000 extreme rows with the LDS row limit raised to
    2^20: its charge sum (about 1.1e19) passes 2^63 inside one LDS table, so
    the carry into the high word must be exact.  Own context, so the knob
    does not leak."""
    n = 100_000
    t = orc.fact_table(n, groups=1)
    t["price"][:] = (1 << 31) - 1
    t["discount"][:] = -128
    t["tax"][:] = 127
    t["date"][:] = 0
    with pimdal_amd.Context(0) as c2:
        c2.set_tune("gb_lds_max_rows", 1 << 20)
        d = {c: c2.to_device(v) for c, v in t.items()}
        got = ops.filter_groupby_q1(c2, d, 2505, by_key=True)
    exp = orc.q1_key(t, 2505)
    assert exp["sum_charge"][0] > (1 << 63)
    _cmp_q1(got, exp, ordered=False)


def test_q1_by_orderkey_few_hot_keys(ctx, orc):
    """Few groups with many rows each: partitions larger than an LDS table's
    row limit go to the HBM tables."""
    n = 1 << 20
    t = orc.fact_table(n, groups=7)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505, by_key=True)
    _cmp_q1(got, orc.q1_key(t, 2505), ordered=False)


def test_hash_aggr_many_partitions(orc):
    """2^16 partitions (beyond the full partition-id histogram's 2^14): the
    per-digit histogram path of the hashed partitioning.  Own context, so
    the knobs do not leak into other tests."""
    import pimdal_amd
    n = 1 << 20
    k = orc.gen_range(np.uint32, 41, 1, n, 0, n // 3)
    v = orc.gen_range(np.int64, 41, 2, n, -(1 << 40), 1 << 40)
    with pimdal_amd.Context(0) as c2:
        c2.set_tune("aggr_dense_max", 0)   # the hashed partitions, not the key-range path
        c2.set_tune("gb_rows_per_part", 16)
        c2.set_tune("gb_groups_hint", n // 3)
        got = agg(c2, ops.hash_aggr, k, v)
    exp = orc.hash_aggr(k, v)
    assert all(np.array_equal(a, b) for a, b in zip(got, exp))


def test_hash_aggr_partitions_over_the_row_limit_between_lds_ones(orc):
    """65536 partitions of about 16 rows; the ones over a 20-row limit (about
    one in eight) go to the HBM tables, the rest stay in LDS tables.  Own
    context, so the knobs do not leak."""
    import pimdal_amd
    n = 1 << 20
    k = orc.gen_range(np.uint32, 43, 1, n, 0, n // 2)
    v = orc.gen_range(np.int64, 43, 2, n, -(1 << 40), 1 << 40)
    with pimdal_amd.Context(0) as c2:
        c2.set_tune("aggr_dense_max", 0)   # the hashed partitions, not the key-range path
        c2.set_tune("gb_rows_per_part", 16)
        c2.set_tune("gb_lds_max_rows", 20)
        c2.set_tune("gb_groups_hint", n // 2)
        got = agg(c2, ops.hash_aggr, k, v)
    exp = orc.hash_aggr(k, v)
    assert all(np.array_equal(a, b) for a, b in zip(got, exp))


def test_q1_by_orderkey_tables_overflowing_between_lds_ones(ctx, orc):
    """2048 partitions with about 768 groups each (the overflow point of a
    1024-slot table): about half overflow to the HBM tables, the rest are
    aggregated in LDS."""
    n = 1 << 22
    t = orc.fact_table(n, groups=2048 * 768)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505, by_key=True)
    _cmp_q1(got, orc.q1_key(t, 2505), ordered=False)


@pytest.mark.parametrize("case", ["late_value", "late_values_past_8_groups", "late_value_filtered_out"])
def test_q1_flags_values_after_the_code_sample(ctx, orc, case):
    """The dense path takes its flag codes from the first 65536 rows; a kept
    row with a flag value the sample lacks sends the call to the codes of
    every row (or, past 8 groups, to the general path); a row the date
    filter drops does not."""
    n = 200_000
    t = orc.fact_table(n)
    if case == "late_value":
        t["flag1"][150_000] = 9
        t["date"][150_000] = 10
    elif case == "late_values_past_8_groups":
        t["flag1"][100_000:100_010] = np.arange(3, 13, dtype=np.uint8)
        t["date"][100_000:100_010] = 10
    else:
        t["flag1"][150_000] = 9
        t["date"][150_000] = 2556
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2505, capacity=4096)
    _cmp_q1(got, orc.q1_flags(t, 2505), ordered=(case != "late_values_past_8_groups"))


@pytest.mark.parametrize("n1,n2", [(2, 2), (4, 2), (1, 1)])
def test_q1_flags_dense_group_counts(ctx, orc, n1, n2):
    """The dense path instantiated for 2, 4 and 8 group codes."""
    n = 300_001
    t = orc.fact_table(n)
    t["flag1"] = (orc.gen_range(np.uint8, 51, 17, n, 0, n1 - 1) * 7 + 65).astype(np.uint8)
    t["flag2"] = (orc.gen_range(np.uint8, 51, 18, n, 0, n2 - 1) * 11 + 70).astype(np.uint8)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    got = ops.filter_groupby_q1(ctx, d, 2000)
    _cmp_q1(got, orc.q1_flags(t, 2000), ordered=True)


@pytest.mark.parametrize("n,groups", [(100_000, 25_000), (1 << 21, 1 << 19)])
def test_q1_by_orderkey_2048_slot_tables(ctx, orc, n, groups):
    """Key group-by with 2048-slot LDS tables (one block per CU)."""
    t = orc.fact_table(n, groups=groups)
    d = {c: ctx.to_device(v) for c, v in t.items()}
    ctx.set_tune("gb_q1_slots", 2048)
    try:
        got = ops.filter_groupby_q1(ctx, d, 2505, by_key=True)
    finally:
        ctx.set_tune("gb_q1_slots", 1024)
    _cmp_q1(got, orc.q1_key(t, 2505), ordered=False)


@pytest.mark.parametrize("n,rows_per_part,bad_row", [((1 << 21) + 4097, 2048, None), (3_000_001, 64, None),
                                                   (700_003, 2048, None), (1_000_003, 512, ("price", 1 << 31))])
def test_q1_by_orderkey_fused_first_partition_pass(ctx, orc, n, rows_per_part, bad_row):
    """Filter + pack fused with the first hash partition pass (partitions of
    9 to 16 bits: 8192-row tiles placed by the low bits, then the stable
    pass over the top 8) and the separate pack + partition passes give the
    oracle's groups, through the synchronous and the device-status forms;
    ragged sizes, 16 partition bits, columns starting one element into
    their allocations, and a kept row beyond the packed widths (general
    path; the device-status form reports unsupported)."""
    t = orc.fact_table(n + 1, groups=max(1, n // 4))
    if bad_row:
        t["date"][4321] = 0
        t[bad_row[0]][4321] = bad_row[1]
    full = {c: ctx.to_device(v) for c, v in t.items()}
    d = {c: v.view(1, n) for c, v in full.items()}          # 4- and 8-byte columns off 16-byte alignment
    h = {c: v[1:] for c, v in t.items()}
    exp = orc.q1_key(h, 2505)
    ctx.set_tune("gb_rows_per_part", rows_per_part)
    try:
        for fused in (1, 0):
            ctx.set_tune("gb_q1_fused", fused)
            _cmp_q1(ops.filter_groupby_q1(ctx, d, 2505, by_key=True), exp, ordered=False)
            out, ng, st = ops.filter_groupby_q1_async(ctx, d, 2505, by_key=True)
            ctx.sync()
            status, got = ops.q1_to_host(out, ng, st)
            if bad_row:
                assert status == -2, fused          # PIMDAL_E_UNSUPPORTED
            else:
                assert status == 0, fused
                got.pop("groups")
                _cmp_q1(got, exp, ordered=False)
    finally:
        ctx.clear_tune("gb_q1_fused")
        ctx.clear_tune("gb_rows_per_part")


# ---------------- TPC-H Q6 shape ----------------

@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 3, 4095, 1 << 20, (1 << 20) + 3])
def test_q6_matches_oracle(ctx, orc, n):
    t = orc.fact_table(n, seed=90 + n)
    d = {c: ctx.to_device(t[c]) for c in ("date", "quantity", "price", "discount")}
    assert ops.q6(ctx, d) == orc.q6(t)
    # other parameters: two years, discount 0..3, quantity < 50
    assert ops.q6(ctx, d, 365, 1096, 0, 3, 50) == orc.q6(t, 365, 1096, 0, 3, 50)


@pytest.mark.gpu
def test_q6_bounds_and_wrapping_sums(ctx, orc):
    date = np.array([730, 731, 1095, 1096, 900, 900, 900, 900, 900, 900] * 7, np.int32)
    disc = np.array([6, 6, 6, 6, 4, 5, 7, 8, 6, 6] * 7, np.int64)
    qty = np.array([10, 10, 10, 10, 10, 10, 10, 10, 23, 24] * 7, np.int64)
    price = np.full(70, (1 << 62) - 5, np.int64)
    t = {"date": date, "quantity": qty, "price": price, "discount": disc}
    d = {c: ctx.to_device(v) for c, v in t.items()}
    assert ops.q6(ctx, d) == orc.q6(t)
    assert ops.q6(ctx, d)[1] == 35


@pytest.mark.gpu
def test_q6_unaligned_columns(ctx, orc):
    # columns starting 4 and 8 bytes past an aligned base take the scalar loads
    n = 100_003
    t = orc.fact_table(n + 1, seed=5)
    d = {c: ctx.to_device(t[c]) for c in ("date", "quantity", "price", "discount")}
    v = {c: d[c].view(1, n) for c in d}
    sub = {c: t[c][1:] for c in d}
    assert ops.q6(ctx, v) == orc.q6(sub)


@pytest.mark.gpu
def test_q6_6e7_rows_full_size(ctx, orc):
    n = 60_000_000
    d = datagen.fact_table(ctx, n)
    t = {c: d[c].numpy() for c in ("date", "quantity", "price", "discount")}
    assert ops.q6(ctx, d) == orc.q6(t)


def _direct_ran(ctx, fn, dk, dv):
    ctx.set_profiling(True)
    ctx.reset_stats()
    try:
        ok, os_, oc, g = fn(ctx, dk, dv)
        return (ok.numpy(g), os_.numpy(g), oc.numpy(g)), ctx.kernel_stat("ad_direct")[0]
    finally:
        ctx.set_profiling(False)


@pytest.mark.parametrize("case", ["k50", "k4000", "span_8192", "top_of_u32", "below_sample_window",
                                  "beyond_sample_window", "beyond_small_window", "few_rows"])
def test_aggr_direct_key_window(ctx, orc, case):
    """The key-indexed path: a sample of the first 2^16 keys places a window
    of 512 to 8192 keys (at least twice the sample's span, one LDS copy per
    wave while 16 fit); keys outside it (met later in the rows) send the
    call to the table paths.  Sort output in its own order, hash canonical."""
    n = (1 << 21) + 3
    v = orc.gen_range(np.int64, 37, 2, n, -(1 << 45), 1 << 45)
    if case == "k50":
        k = orc.gen_range(np.uint32, 37, 1, n, 1, 50)
    elif case == "k4000":
        k = orc.gen_range(np.uint32, 37, 1, n, 1, 4000)
    elif case == "span_8192":
        k = orc.gen_range(np.uint32, 37, 1, n, 100, 100 + 8191)
        k[0], k[1] = 100, 100 + 8191
    elif case == "top_of_u32":
        k = orc.gen_range(np.uint32, 37, 1, n, 0xFFFFFFFF - 3000, 0xFFFFFFFF)
        k[::5] = 0xFFFFFFFF
    elif case == "below_sample_window":
        k = orc.gen_range(np.uint32, 37, 1, n, 20_000, 20_100)
        k[-7] = 5                       # past the sample, below the window
    elif case == "beyond_sample_window":
        k = orc.gen_range(np.uint32, 37, 1, n, 0, 100)
        k[(1 << 16) + 5] = 50_000       # past the sample, above the window
    elif case == "beyond_small_window":
        k = orc.gen_range(np.uint32, 37, 1, n, 1000, 1100)
        k[(1 << 16) + 9] = 1000 + 700   # within 8192 of the sample, past its 512-key window
    else:
        n = 1000
        k = orc.gen_range(np.uint32, 37, 1, n, 7, 300)
        v = v[:n]
    exp = orc.sort_aggr(k, v)
    dk, dv = ctx.to_device(k), ctx.to_device(v)
    for fn in (ops.hash_aggr, ops.sort_aggr):
        got, launches = _direct_ran(ctx, fn, dk, dv)
        assert launches >= 1, (case, fn.__name__)
        if fn is ops.hash_aggr:
            got = canon(*got)
        assert all(np.array_equal(a, b) for a, b in zip(got, exp)), (case, fn.__name__)
