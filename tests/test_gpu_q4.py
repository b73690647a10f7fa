"""TPC-H Q4 shape on the GPU (qualifying orders -> orders table ->
lineitem pass marking the orders that have a late lineitem -> counts per
priority) against the CPU restatement, bit-exact."""
import numpy as np
import pytest

from pimdal_amd import _lib, datagen, operators as ops

pytestmark = pytest.mark.gpu

LO, HI = (int(x) for x in datagen.pack_days(np.array([8582, 8674])))   # 1993-07-01, 1993-10-01


def _dev(ctx, t):
    return {k: ctx.to_device(v) for k, v in t.items()}


def _check(ctx, orc, t, lo, hi, capacity=256):
    codes, counts = orc.q4(t, lo, hi)
    gp, gc, g = ops.tpch_q4(ctx, _dev(ctx, t), lo, hi, capacity=capacity)
    assert g == codes.size
    k = min(g, capacity)
    assert np.array_equal(gp, codes[:k]) and np.array_equal(gc, counts[:k])
    return g


@pytest.mark.parametrize("n", [40, 1000, 1 << 20])
def test_q4_matches_oracle(ctx, orc, n):
    t = datagen.q4_tables(n, seed=3 + n)
    _check(ctx, orc, t, LO, HI)


@pytest.mark.parametrize("days", [(8035, 10800), (9000, 9001), (9000, 8000), (20000, 21000)])
def test_q4_other_date_ranges(ctx, orc, days):
    """Every order, one day, an empty range and a range with no orders."""
    t = datagen.q4_tables(300_000, seed=23)
    lo, hi = (int(x) for x in datagen.pack_days(np.array(days)))
    _check(ctx, orc, t, lo, hi)


def test_q4_edge_cases(ctx, orc):
    """Orderkeys at the int32 extremes, lineitems of missing orders, equal
    commit and receipt dates (not late), priority codes up to 255, and a
    capacity below the group count."""
    t = datagen.q4_tables(50_000, seed=43)
    t["o_orderkey"][0] = np.int32(-(1 << 31))
    t["o_orderkey"][1] = np.int32((1 << 31) - 1)
    t["o_orderdate"][:2] = LO
    t["l_orderkey"][:50] = t["o_orderkey"][0]
    t["l_orderkey"][50:100] = t["o_orderkey"][1]
    t["l_orderkey"][100:150] = 77_777_777          # no such order
    t["l_receiptdate"][::5] = t["l_commitdate"][::5]
    t["o_orderpriority"][::3] = 255
    t["o_orderpriority"][1::7] = 0
    g = _check(ctx, orc, t, LO, HI)
    assert g == 6
    _check(ctx, orc, t, LO, HI, capacity=2)


def test_q4_sf10_lineitems(ctx, orc):
    """SF10's shape: 6e7 lineitems and 1.5e7 orders, every priority's count."""
    t = datagen.q4_tables(60_000_000, seed=33)
    g = _check(ctx, orc, t, LO, HI)
    assert g == 5


def test_q4_repeated_orderkey_is_refused(ctx):
    t = datagen.q4_tables(20_000, seed=53)
    t["o_orderkey"][:] = t["o_orderkey"][0]
    t["o_orderdate"][:] = LO
    with pytest.raises(_lib.PimdalError, match="unique"):
        ops.tpch_q4(ctx, _dev(ctx, t), LO, HI)
