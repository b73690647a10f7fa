"""TPC-H Q3 shape on the GPU (customer set -> orders table -> lineitem probe
with the group-by in the table -> TPC-H order) against the CPU restatement:
every group, in order, bit-exact."""
import ctypes as C

import numpy as np
import pytest

from pimdal_amd import _lib, datagen, operators as ops

pytestmark = pytest.mark.gpu

DATE = int(datagen.pack_days(np.array([9204]))[0])   # 1995-03-15
BUILDING = datagen.Q3_SEGMENTS.index("BUILDING")


def _dev(ctx, t):
    return {k: ctx.to_device(v) for k, v in t.items()}


def _check(got, exp):
    for f in ("key", "revenue", "orderdate", "shippriority"):
        assert np.array_equal(got[f], exp[f]), f


@pytest.mark.parametrize("n", [40, 1000, 1 << 20])
def test_q3_all_groups_match_oracle(ctx, orc, n):
    t = datagen.q3_tables(n, seed=11 + n)
    exp, g_exp = orc.q3(t, BUILDING, DATE)
    got, g = ops.tpch_q3(ctx, _dev(ctx, t), BUILDING, DATE, capacity=max(g_exp, 1))
    assert g == g_exp
    _check(got, exp)


@pytest.mark.parametrize("segment,day", [(0, 8035 + 600), (4, 8035 + 2000), (2, 8035)])
def test_q3_other_segments_and_dates(ctx, orc, segment, day):
    t = datagen.q3_tables(300_000, seed=21)
    date = int(datagen.pack_days(np.array([day]))[0])
    exp, g_exp = orc.q3(t, segment, date)
    got, g = ops.tpch_q3(ctx, _dev(ctx, t), segment, date, capacity=max(g_exp, 1))
    assert g == g_exp
    _check(got, exp)


def test_q3_top10_at_6e7_lineitems(ctx, orc):
    """SF10's shape: 6e7 lineitems, 1.5e7 orders, 1.5e6 customers; every
    group compared in order, and the top 10 as the reference reports them."""
    t = datagen.q3_tables(60_000_000, seed=31)
    exp, g_exp = orc.q3(t, BUILDING, DATE)
    d = _dev(ctx, t)
    got, g = ops.tpch_q3(ctx, d, BUILDING, DATE, capacity=g_exp)
    assert g == g_exp and g > 10_000
    _check(got, exp)
    top, g10 = ops.tpch_q3(ctx, d, BUILDING, DATE, capacity=10)
    assert g10 == g_exp
    _check(top, {f: v[:10] for f, v in exp.items()})


def test_q3_edge_cases(ctx, orc):
    """Negative prices (revenue below zero sorts last), keys at the int32
    extremes, lineitems whose order is missing or filtered, a segment with no
    customers."""
    t = datagen.q3_tables(50_000, seed=41)
    t["l_extendedprice"][::7] *= -1
    t["o_orderkey"][0] = np.int32(-(1 << 31))
    t["o_orderkey"][1] = np.int32((1 << 31) - 1)
    t["l_orderkey"][:50] = t["o_orderkey"][0]
    t["l_orderkey"][50:100] = t["o_orderkey"][1]
    t["l_orderkey"][100:150] = 77_777_777          # no such order
    for seg in (BUILDING, 9):
        exp, g_exp = orc.q3(t, seg, DATE)
        got, g = ops.tpch_q3(ctx, _dev(ctx, t), seg, DATE, capacity=max(g_exp, 1))
        assert g == g_exp
        _check(got, exp)


def test_q3_custkeys_beyond_the_bitmap(ctx, orc):
    """Customer keys below the bitmap's size (the power of two above the
    customer count, 2^16 here) go to the bitmap, the others (up to 2^31 - 1,
    and negative ones, 2^31 and up as unsigned) to the hash set; keys on
    both sides of the boundary."""
    t = datagen.q3_tables(200_000, seed=61)
    old = t["c_custkey"].astype(np.int64)
    new = np.where(old % 3 == 0, old + (1 << 25) + 10, np.where(old % 3 == 1, -old, old))
    new[:6] = [65535, 65536, (1 << 25) - 1, 1 << 25, (1 << 31) - 1, -(1 << 31)]
    assert np.unique(new).size == new.size
    remap = np.zeros(old.max() + 1, np.int64)
    remap[old] = new
    t["c_custkey"] = new.astype(np.int32)
    t["o_custkey"] = remap[t["o_custkey"]].astype(np.int32)
    for seg in range(5):
        exp, g_exp = orc.q3(t, seg, DATE)
        got, g = ops.tpch_q3(ctx, _dev(ctx, t), seg, DATE, capacity=max(g_exp, 1))
        assert g == g_exp and g > 0
        _check(got, exp)


def test_q3_repeated_orderkey_is_refused(ctx):
    t = datagen.q3_tables(20_000, seed=51)
    t["o_orderkey"][:] = t["o_orderkey"][0]
    t["o_orderdate"][:] = t["o_orderdate"].min()
    with pytest.raises(_lib.PimdalError, match="unique"):
        ops.tpch_q3(ctx, _dev(ctx, t), BUILDING, DATE)


@pytest.mark.parametrize("cap", [1, 2, 10, 100, 4096, 4097])
def test_q3_first_groups_match_oracle(ctx, orc, cap):
    """A capacity below the group count: the revenue histogram finds the
    candidates of the first `cap` groups, sorted in one block (up to 4096);
    4097 takes the full sort."""
    t = datagen.q3_tables(1 << 21, seed=61)
    exp, g_exp = orc.q3(t, BUILDING, DATE, cap=cap)
    got, g = ops.tpch_q3(ctx, _dev(ctx, t), BUILDING, DATE, capacity=cap)
    assert g == g_exp and g > cap
    _check(got, exp)


def test_q3_first_groups_with_tied_and_negative_revenue(ctx, orc):
    """Revenues tied over thousands of groups (more candidates than one block
    sorts: the full sort) and a few negative ones; ties ordered by date,
    then orderkey."""
    t = datagen.q3_tables(1 << 20, seed=62)
    t["l_extendedprice"][:] = 1000
    t["l_discount"][:] = 0
    t["l_extendedprice"][::97] = -5
    for cap in (10, 3000):
        exp, g_exp = orc.q3(t, BUILDING, DATE, cap=cap)
        got, g = ops.tpch_q3(ctx, _dev(ctx, t), BUILDING, DATE, capacity=cap)
        assert g == g_exp
        _check(got, exp)
