"""Synthetic workloads generated on the device (counter-based, seed-fixed).

Each recipe produces the columns of one benchmark workload directly in HBM
with the generators of include/pimdal_gen.h, so the same (seed, column, row)
gives the same value as the host restatement.
"""
from __future__ import annotations

import numpy as np

from ._lib import call
from .device import Context, DeviceArray

COL = dict(KEY=1, PAYLOAD=2, BUILD_PERM=3, PROBE_FK=4, BUILD_PAY=5, PROBE_PAY=6, ZIPF=7, QTY=12, PRICE=13,
           DISC=14, TAX=15, DATE=16, FLAG1=17, FLAG2=18)

SEEDS = {"select": 0x5EED0001, "sort": 0x5EED0002, "join": 0x5EED0003, "groupby": 0x5EED0004, "serve": 0x5EED0005,
         "aggr": 0x5EED0007}

SELECT_THRESHOLD = 429496730  # ceil(0.1 * 2^32)


def u32(ctx: Context, seed: int, col: int, n: int) -> DeviceArray:
    d = ctx.empty(n, np.uint32)
    call("pimdal_gen_u32", ctx.handle, seed, col, n, d.vp)
    return d


def rng_range(ctx: Context, dtype, seed: int, col: int, n: int, lo: int, hi: int) -> DeviceArray:
    from .device import pimdal_type
    d = ctx.empty(n, dtype)
    call("pimdal_gen_range", ctx.handle, pimdal_type(dtype), seed, col, n, lo, hi, d.vp)
    return d


def iota(ctx: Context, n: int) -> DeviceArray:
    """Row ids 0..n-1 (the identity permutation restricted to n)."""
    d = ctx.empty(n, np.uint32)
    call("pimdal_gen_iota", ctx.handle, n, 0, d.vp)
    return d


def perm(ctx: Context, seed: int, col: int, m: int, n: int, offset: int = 0) -> DeviceArray:
    d = ctx.empty(n, np.uint32)
    call("pimdal_gen_perm", ctx.handle, seed, col, m, offset, n, d.vp)
    return d


def zipf_ranks(ctx: Context, seed: int, col: int, m: int, s: float, n: int) -> DeviceArray:
    d = ctx.empty(n, np.uint32)
    call("pimdal_gen_zipf", ctx.handle, seed, col, m, float(s), n, d.vp)
    return d


# ---------------- workload recipes ----------------

def select_workload(ctx: Context, n: int, seed: int = SEEDS["select"]):
    """uint32 key uniform on [0, 2^32) and uint32 payload."""
    return u32(ctx, seed, COL["KEY"], n), u32(ctx, seed, COL["PAYLOAD"], n)


def sort_workload(ctx: Context, n: int, zipf: bool = False, seed: int = SEEDS["sort"], distinct: int = 1 << 24):
    """Keys uniform u32, or fmix32(Zipf(1.0) rank over `distinct` values); payload = row id."""
    if zipf:
        r = zipf_ranks(ctx, seed, COL["ZIPF"], distinct, 1.0, n)
        k = ctx.empty(n, np.uint32)
        call("pimdal_fmix32", ctx.handle, n, r.vp, k.vp)
        r.free()
    else:
        k = u32(ctx, seed, COL["KEY"], n)
    return k, iota(ctx, n)


def join_workload(ctx: Context, nr: int, ns: int, zipf: bool = False, seed: int = SEEDS["join"]):
    """Build: unique keys perm(nr); probe: FK uniform over [0, nr) or Zipf(1.0)
    ranks mapped through the build permutation; int32 payloads."""
    rk = perm(ctx, seed, COL["BUILD_PERM"], nr, nr)
    rp = u32(ctx, seed, COL["BUILD_PAY"], nr)
    if zipf:
        ranks = zipf_ranks(ctx, seed, COL["ZIPF"], nr, 1.0, ns)
        # rank r in [1, nr] -> build key perm(r - 1)
        sk = ctx.empty(ns, np.uint32)
        call("pimdal_perm_apply", ctx.handle, seed, COL["BUILD_PERM"], nr, ns, -1, ranks.vp, sk.vp)
        ranks.free()
    else:
        sk = rng_range(ctx, np.uint32, seed, COL["PROBE_FK"], ns, 0, nr - 1)
    sp = u32(ctx, seed, COL["PROBE_PAY"], ns)
    return rk, rp, sk, sp


def aggr_workload(ctx: Context, n: int, lo: int, hi: int, seed: int = SEEDS["aggr"]):
    """Group-by input: uint32 key uniform on [lo, hi], int64 value on [1, 10]
    (reference/db_operators/aggregate.cpp:33-36 draws keys on [1, 50] and
    values on [1, 10]; the PIM microbenchmarks use keys on [1, 4000] and
    [1, 25])."""
    return (rng_range(ctx, np.uint32, seed, COL["KEY"], n, lo, hi),
            rng_range(ctx, np.int64, seed, COL["PAYLOAD"], n, 1, 10))


def fact_table(ctx: Context, n: int, seed: int = SEEDS["groupby"], groups: int | None = None):
    """The config-4 fact table: dict of device columns."""
    if groups is None:
        groups = max(1, n // 4)
    ok = ctx.empty(n, np.int32)
    call("pimdal_gen_orderkey", ctx.handle, seed, n, groups, ok.vp)
    return {
        "orderkey": ok,
        "quantity": rng_range(ctx, np.int64, seed, COL["QTY"], n, 1, 50),
        "price": rng_range(ctx, np.int64, seed, COL["PRICE"], n, 90000, 10500000),
        "discount": rng_range(ctx, np.int64, seed, COL["DISC"], n, 0, 10),
        "tax": rng_range(ctx, np.int64, seed, COL["TAX"], n, 0, 8),
        "date": rng_range(ctx, np.int32, seed, COL["DATE"], n, 0, 2556),
        "flag1": rng_range(ctx, np.uint8, seed, COL["FLAG1"], n, 0, 2),
        "flag2": rng_range(ctx, np.uint8, seed, COL["FLAG2"], n, 0, 1),
    }


def dim_table(ctx: Context, n_dim: int, seed: int = SEEDS["groupby"]):
    """Dimension table: unique int32 key (the fact orderkey values, each once),
    int32 date on [0, 2557), uint8 category on [0, 25)."""
    k = ctx.empty(n_dim, np.int32)
    call("pimdal_gen_dim_keys", ctx.handle, seed, n_dim, k.vp)
    return {"key": k, "date": rng_range(ctx, np.int32, seed, 20, n_dim, 0, 2556),
            "category": rng_range(ctx, np.uint8, seed, 21, n_dim, 0, 24)}


# ---------------- TPC-H Q3 shape (host-generated, seeded) ----------------

Q3_SEGMENTS = ("AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD", "MACHINERY")   # dictionary order


def pack_days(days: np.ndarray) -> np.ndarray:
    """Days since 1970-01-01 to the packed date day + (month << 5) +
    ((year - 1900) << 9) (pimdal/main/tpc_h/reader/read_table.cpp:8-16)."""
    d = np.datetime64("1970-01-01", "D") + np.asarray(days, dtype=np.int64).astype("timedelta64[D]")
    y = d.astype("datetime64[Y]").astype(np.int64) + 1970
    m = d.astype("datetime64[M]").astype(np.int64) % 12 + 1
    day = (d - d.astype("datetime64[M]")).astype(np.int64) + 1
    return (day + (m << 5) + ((y - 1900) << 9)).astype(np.int32)


def q3_tables(n_lineitem: int, seed: int = 0x5EED0006) -> dict:
    """Synthetic customer / orders / lineitem columns in TPC-H's shape: 40
    lineitems per customer and 4 per order on average (SF10 at 6e7
    lineitems); sparse unique orderkeys; order dates uniform over
    1992-01-01 .. 1998-08-02; each lineitem ships 1..121 days after its order
    (lineitems grouped by order); prices in cents, discount 0..10 percent;
    segment codes index Q3_SEGMENTS.  Dates packed."""
    rng = np.random.default_rng(seed)
    nl = int(n_lineitem)
    no = max(1, nl // 4)
    nc = max(1, nl // 40)
    start = 8035                      # 1992-01-01
    span = 2405 - 151 + 1             # order dates up to 1998-08-02
    c_key = (rng.permutation(nc) + 1).astype(np.int32)
    c_seg = rng.integers(0, len(Q3_SEGMENTS), nc).astype(np.uint8)
    o_key = (np.arange(no, dtype=np.int64) * 4 + rng.integers(0, 4, no)).astype(np.int32)
    o_key = o_key[rng.permutation(no)]
    o_cust = c_key[rng.integers(0, nc, no)]
    o_days = start + rng.integers(0, span, no)
    o_prio = rng.integers(0, 2, no).astype(np.int32)
    l_orow = np.sort(rng.integers(0, no, nl))
    l_days = o_days[l_orow] + rng.integers(1, 122, nl)
    return {
        "c_custkey": c_key, "c_mktsegment": c_seg,
        "o_orderkey": o_key, "o_custkey": o_cust, "o_orderdate": pack_days(o_days), "o_shippriority": o_prio,
        "l_orderkey": o_key[l_orow], "l_extendedprice": rng.integers(90000, 10500001, nl).astype(np.int64),
        "l_discount": rng.integers(0, 11, nl).astype(np.int64), "l_shipdate": pack_days(l_days),
    }


Q4_PRIORITIES = ("1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED", "5-LOW")   # dictionary order


def q4_tables(n_lineitem: int, seed: int = 0x5EED0007) -> dict:
    """Synthetic orders / lineitem columns in TPC-H's shape for Q4: 4
    lineitems per order on average (SF10 at 6e7 lineitems), sparse unique
    orderkeys, order dates uniform over 1992-01-01 .. 1998-08-02, commit
    dates 30..90 days after the order, receipt dates 1..30 days after a ship
    date 1..121 days after the order (dbgen's ranges), priority codes
    indexing Q4_PRIORITIES.  Dates packed."""
    rng = np.random.default_rng(seed)
    nl = int(n_lineitem)
    no = max(1, nl // 4)
    start = 8035                      # 1992-01-01
    span = 2405 - 151 + 1
    o_key = (np.arange(no, dtype=np.int64) * 4 + rng.integers(0, 4, no)).astype(np.int32)
    o_key = o_key[rng.permutation(no)]
    o_days = start + rng.integers(0, span, no)
    l_orow = np.sort(rng.integers(0, no, nl))
    ship = o_days[l_orow] + rng.integers(1, 122, nl)
    return {
        "o_orderkey": o_key, "o_orderdate": pack_days(o_days),
        "o_orderpriority": rng.integers(0, len(Q4_PRIORITIES), no).astype(np.uint8),
        "l_orderkey": o_key[l_orow], "l_commitdate": pack_days(o_days[l_orow] + rng.integers(30, 91, nl)),
        "l_receiptdate": pack_days(ship + rng.integers(1, 31, nl)),
    }


Q5_REGIONS = ("AFRICA", "AMERICA", "ASIA", "EUROPE", "MIDDLE EAST")   # dictionary order
# TPC-H's 25 nations by key and their region keys
Q5_NATION_REGION = (0, 1, 1, 1, 4, 0, 3, 3, 2, 2, 4, 4, 2, 4, 0, 0, 0, 1, 2, 3, 4, 2, 3, 3, 1)


def q5_tables(n_lineitem: int, seed: int = 0x5EED0008) -> dict:
    """Synthetic region / nation / customer / orders / lineitem / supplier
    columns in TPC-H's shape for Q5: 25 nations in 5 regions (dbgen's
    mapping), 40 lineitems per customer, 4 per order, 600 per supplier
    (SF10 at 6e7 lineitems: 1.5e6 customers, 1e5 suppliers); order dates
    uniform over 1992-01-01 .. 1998-08-02; prices in cents, discount 0..10
    percent; region names as codes into Q5_REGIONS.  Dates packed."""
    rng = np.random.default_rng(seed)
    nl = int(n_lineitem)
    no, nc, ns = max(1, nl // 4), max(1, nl // 40), max(1, nl // 600)
    o_key = (np.arange(no, dtype=np.int64) * 4 + rng.integers(0, 4, no)).astype(np.int32)
    o_key = o_key[rng.permutation(no)]
    c_key = (rng.permutation(nc) + 1).astype(np.int32)
    s_key = (rng.permutation(ns) + 1).astype(np.int32)
    o_days = 8035 + rng.integers(0, 2405 - 151 + 1, no)
    l_orow = np.sort(rng.integers(0, no, nl))
    return {
        "r_regionkey": np.arange(5, dtype=np.int32), "r_name": np.arange(5, dtype=np.uint8),
        "n_nationkey": np.arange(25, dtype=np.int32), "n_regionkey": np.array(Q5_NATION_REGION, np.int32),
        "c_custkey": c_key, "c_nationkey": rng.integers(0, 25, nc).astype(np.int32),
        "o_orderkey": o_key, "o_custkey": c_key[rng.integers(0, nc, no)], "o_orderdate": pack_days(o_days),
        "l_orderkey": o_key[l_orow], "l_suppkey": s_key[rng.integers(0, ns, nl)],
        "l_extendedprice": rng.integers(90000, 10500001, nl).astype(np.int64),
        "l_discount": rng.integers(0, 11, nl).astype(np.int64),
        "s_suppkey": s_key, "s_nationkey": rng.integers(0, 25, ns).astype(np.int32),
    }
