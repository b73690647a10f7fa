"""TPC-H Q5 shape on the GPU (region's nations -> customers -> orders of the
year -> lineitems -> suppliers of the same nation -> revenue per nation)
against the CPU restatement, bit-exact and in order."""
import numpy as np
import pytest

from pimdal_amd import _lib, datagen, operators as ops

pytestmark = pytest.mark.gpu

ASIA = datagen.Q5_REGIONS.index("ASIA")
LO, HI = (int(x) for x in datagen.pack_days(np.array([8766, 9131])))   # 1994-01-01, 1995-01-01


def _dev(ctx, t):
    return {k: ctx.to_device(v) for k, v in t.items()}


def _check(ctx, orc, t, name, lo, hi, capacity=1024):
    en, er = orc.q5(t, name, lo, hi)
    gn, gr, g = ops.tpch_q5(ctx, _dev(ctx, t), name, lo, hi, capacity=capacity)
    assert g == en.size
    k = min(g, capacity)
    assert np.array_equal(gn, en[:k]) and np.array_equal(gr, er[:k])
    return g


@pytest.mark.parametrize("n", [600, 20_000, 1 << 20])
def test_q5_matches_oracle(ctx, orc, n):
    t = datagen.q5_tables(n, seed=5 + n)
    _check(ctx, orc, t, ASIA, LO, HI)


@pytest.mark.parametrize("region", range(5))
def test_q5_every_region_and_date_ranges(ctx, orc, region):
    t = datagen.q5_tables(300_000, seed=25)
    _check(ctx, orc, t, region, LO, HI)
    lo2, hi2 = (int(x) for x in datagen.pack_days(np.array([8035, 10800])))
    _check(ctx, orc, t, region, lo2, hi2)
    _check(ctx, orc, t, region, HI, LO)          # empty range


def test_q5_edge_cases(ctx, orc):
    """Negative prices, nation keys off the nation table, customers and
    suppliers outside every region, lineitems of missing orders and
    suppliers, a region name no row has, and a short capacity."""
    t = datagen.q5_tables(60_000, seed=45)
    t["l_extendedprice"][::7] *= -1
    t["c_nationkey"][::11] = 1000                 # in range, not a nation
    t["s_nationkey"][::13] = -3                   # out of range: matches nothing
    t["l_orderkey"][:100] = 77_777_777
    t["l_suppkey"][100:200] = 88_888_888
    g = _check(ctx, orc, t, ASIA, LO, HI)
    assert g == 5
    _check(ctx, orc, t, ASIA, LO, HI, capacity=2)
    _check(ctx, orc, t, 9, LO, HI)


def test_q5_sf10_lineitems(ctx, orc):
    """SF10's shape: 6e7 lineitems, 1.5e7 orders, 1.5e6 customers, 1e5 suppliers."""
    t = datagen.q5_tables(60_000_000, seed=35)
    assert _check(ctx, orc, t, ASIA, LO, HI) == 5


def test_q5_custkeys_beyond_the_direct_array(ctx, orc):
    """Customer keys below the direct array's size (the power of two above
    the customer count, 2^16 here) are looked up there, larger and negative
    keys in the hash table; keys on both sides of the boundary."""
    t = datagen.q5_tables(100_000, seed=65)
    old = t["c_custkey"].astype(np.int64)
    new = np.where(old % 3 == 0, old + (1 << 25) + 10, np.where(old % 3 == 1, -old, old))
    new[:4] = [65535, 65536, (1 << 31) - 1, -(1 << 31)]
    assert np.unique(new).size == new.size
    remap = np.zeros(old.max() + 1, np.int64)
    remap[old] = new
    t["c_custkey"] = new.astype(np.int32)
    t["o_custkey"] = remap[t["o_custkey"]].astype(np.int32)
    for region in range(5):
        assert _check(ctx, orc, t, region, LO, HI) > 0


def test_q5_refusals(ctx):
    t = datagen.q5_tables(20_000, seed=55)
    bad = dict(t)
    bad["n_nationkey"] = t["n_nationkey"].copy()
    bad["n_nationkey"][3] = 5000
    with pytest.raises(_lib.PimdalError, match="nation keys"):
        ops.tpch_q5(ctx, _dev(ctx, bad), ASIA, LO, HI)
    bad = dict(t)
    bad["s_suppkey"] = np.full_like(t["s_suppkey"], t["s_suppkey"][0])
    with pytest.raises(_lib.PimdalError, match="unique"):
        ops.tpch_q5(ctx, _dev(ctx, bad), ASIA, LO, HI)
    # a region's customer key twice: in the direct array, and in the hash table
    for key in (7, 1 << 30):
        bad = dict(t)
        bad["c_custkey"] = t["c_custkey"].copy()
        bad["c_nationkey"] = t["c_nationkey"].copy()
        bad["c_custkey"][:2] = key
        rk = t["r_regionkey"][t["r_name"] == ASIA][0]
        bad["c_nationkey"][:2] = t["n_nationkey"][t["n_regionkey"] == rk][0]
        with pytest.raises(_lib.PimdalError, match="c_custkey"):
            ops.tpch_q5(ctx, _dev(ctx, bad), ASIA, LO, HI)
