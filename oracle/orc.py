"""ctypes binding of oracle/liboracle.so (the CPU restatement).

TEST INFRASTRUCTURE ONLY: the checker the HIP operators are compared
against, and the CPU baseline timed beside them.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

u64, i64, u32, i32, vp, dbl = C.c_uint64, C.c_int64, C.c_uint32, C.c_int32, C.c_void_p, C.c_double

_SIG = {
    "orc_gen_u32": (None, [u64, u32, u64, vp]),
    "orc_gen_range_i64": (None, [u64, u32, u64, i64, i64, vp]),
    "orc_gen_range_i32": (None, [u64, u32, u64, i64, i64, vp]),
    "orc_gen_range_u8": (None, [u64, u32, u64, i64, i64, vp]),
    "orc_gen_perm": (None, [u64, u32, u64, u64, u64, vp]),
    "orc_perm_apply": (None, [u64, u32, u64, u64, vp, vp]),
    "orc_zipf_cdf": (None, [u64, dbl, vp]),
    "orc_gen_zipf": (None, [u64, u32, vp, u64, u64, vp]),
    "orc_fmix32_array": (None, [u64, vp, vp]),
    "orc_gen_orderkey": (None, [u64, u64, u64, vp]),
    "orc_hash32": (u32, [u32]),
    "orc_select_u32": (u64, [vp, vp, u64, u32, u32, vp, vp]),
    "orc_select_i32": (u64, [vp, vp, u64, i32, i32, vp, vp]),
    "orc_select_i64": (u64, [vp, vp, u64, i64, i64, vp, vp]),
    "orc_sort_pairs_u32": (None, [vp, vp, u64, vp, vp]),
    "orc_sort_pairs_u32_u64": (None, [vp, vp, u64, vp, vp]),
    "orc_sort_keys_u32": (None, [vp, u64, vp]),
    "orc_hash_join_u32": (u64, [vp, vp, u64, vp, vp, u64, u32, vp, vp, u64]),
    "orc_sort_merge_join_u32": (u64, [vp, vp, u64, vp, vp, u64, vp, vp, u64]),
    "orc_hash_aggr_u32_i64": (u64, [vp, vp, u64, vp, vp, vp, u64]),
    "orc_sort_aggr_u32_i64": (u64, [vp, vp, u64, vp, vp, vp, u64]),
    "orc_q1_flags": (u64, [vp, vp, vp, vp, vp, vp, vp, u64, i32, u64, vp, vp, vp, vp, vp, vp, vp]),
    "orc_q1_key": (u64, [vp, vp, vp, vp, vp, vp, u64, i32, u64, vp, vp, vp, vp, vp, vp, vp]),
    "orc_q6": (u64, [vp, vp, vp, vp, u64, i32, i32, i64, i64, i64, vp]),
    "orc_date_pack_str": (C.c_int, [C.c_char_p, vp]),
    "orc_date32_pack": (None, [vp, u64, vp]),
    "orc_bloom_build": (None, [vp, u64, u32, u32, vp]),
    "orc_bloom_probe": (u64, [vp, vp, u64, u32, u32, vp, vp, vp]),
    "orc_pim_hash": (u32, [C.c_int, u32]),
    "orc_q3": (i64, [vp, vp, u64, vp, vp, vp, vp, u64, vp, vp, vp, vp, u64, u32, i32, u64, vp, vp, vp, vp]),
    "orc_q4": (i64, [vp, vp, vp, u64, vp, vp, vp, u64, i32, i32, vp]),
    "orc_q5": (i64, [vp, vp, u64, vp, vp, u64, vp, vp, u64, vp, vp, vp, u64, vp, vp, vp, vp, u64, vp, vp, u64, u32,
                     i32, i32, u64, vp, vp]),
    "orc_num_threads": (C.c_int, []),
    "orc_set_threads": (None, [C.c_int]),
}

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------- generators ----------------

def gen_u32(seed, col, n):
    out = np.empty(n, np.uint32)
    lib().orc_gen_u32(seed, col, n, _p(out))
    return out


def gen_range(dtype, seed, col, n, lo, hi):
    dt = np.dtype(dtype)
    out = np.empty(n, dt)
    fn = {np.dtype(np.int64): "orc_gen_range_i64", np.dtype(np.int32): "orc_gen_range_i32",
          np.dtype(np.uint8): "orc_gen_range_u8"}.get(dt)
    if fn is None:
        if dt == np.dtype(np.uint32):
            tmp = gen_range(np.int64, seed, col, n, lo, hi)
            return tmp.astype(np.uint32)
        raise TypeError(dt)
    getattr(lib(), fn)(seed, col, n, lo, hi, _p(out))
    return out


def gen_perm(seed, col, m, n, offset=0):
    out = np.empty(n, np.uint32)
    lib().orc_gen_perm(seed, col, m, offset, n, _p(out))
    return out


def perm_apply(seed, col, m, a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.empty_like(a)
    lib().orc_perm_apply(seed, col, m, a.size, _p(a), _p(out))
    return out


def zipf_cdf(m, s):
    out = np.empty(m, np.float64)
    lib().orc_zipf_cdf(m, s, _p(out))
    return out


def gen_zipf(seed, col, m, s, n, cdf=None):
    cdf = zipf_cdf(m, s) if cdf is None else cdf
    out = np.empty(n, np.uint32)
    lib().orc_gen_zipf(seed, col, _p(cdf), m, n, _p(out))
    return out


def fmix32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.empty_like(a)
    lib().orc_fmix32_array(a.size, _p(a), _p(out))
    return out


def gen_orderkey(seed, n, groups):
    out = np.empty(n, np.int32)
    lib().orc_gen_orderkey(seed, n, groups, _p(out))
    return out


def hash32(x: int) -> int:
    return lib().orc_hash32(x)


# ---------------- operators ----------------

def select(key, pay, lo, hi):
    key = np.ascontiguousarray(key)
    pay = np.ascontiguousarray(pay, dtype=np.uint32)
    n = key.size
    op = np.empty(n, np.uint32)
    orow = np.empty(n, np.uint32)
    if key.dtype == np.uint32:
        k = lib().orc_select_u32(_p(key), _p(pay), n, lo, hi, _p(op), _p(orow))
    elif key.dtype == np.int32:
        k = lib().orc_select_i32(_p(key), _p(pay), n, lo, hi, _p(op), _p(orow))
    elif key.dtype == np.int64:
        k = lib().orc_select_i64(_p(key), _p(pay), n, lo, hi, _p(op), _p(orow))
    else:
        raise TypeError(key.dtype)
    return op[:k], orow[:k]


def sort_pairs(key, val=None):
    key = np.ascontiguousarray(key, dtype=np.uint32)
    n = key.size
    ok = np.empty(n, np.uint32)
    if val is None:
        lib().orc_sort_pairs_u32(_p(key), None, n, _p(ok), None)
        return ok, None
    val = np.ascontiguousarray(val)
    ov = np.empty_like(val)
    if val.dtype.itemsize == 4:
        lib().orc_sort_pairs_u32(_p(key), _p(val), n, _p(ok), _p(ov))
    else:
        lib().orc_sort_pairs_u32_u64(_p(key), _p(val), n, _p(ok), _p(ov))
    return ok, ov


def sort_keys(key):
    key = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty_like(key)
    lib().orc_sort_keys_u32(_p(key), key.size, _p(out))
    return out


def hash_join(bkey, bpay, pkey, ppay, part_n=16384):
    bkey, bpay, pkey, ppay = (np.ascontiguousarray(a, dtype=np.uint32) for a in (bkey, bpay, pkey, ppay))
    k = lib().orc_hash_join_u32(_p(bkey), _p(bpay), bkey.size, _p(pkey), _p(ppay), pkey.size, part_n, None, None, 0)
    op = np.empty(k, np.uint32)
    ob = np.empty(k, np.uint32)
    lib().orc_hash_join_u32(_p(bkey), _p(bpay), bkey.size, _p(pkey), _p(ppay), pkey.size, part_n, _p(op), _p(ob), k)
    return op, ob


def sort_merge_join(bkey, bpay, pkey, ppay):
    bkey, bpay, pkey, ppay = (np.ascontiguousarray(a, dtype=np.uint32) for a in (bkey, bpay, pkey, ppay))
    k = lib().orc_sort_merge_join_u32(_p(bkey), _p(bpay), bkey.size, _p(pkey), _p(ppay), pkey.size, None, None, 0)
    op = np.empty(k, np.uint32)
    ob = np.empty(k, np.uint32)
    lib().orc_sort_merge_join_u32(_p(bkey), _p(bpay), bkey.size, _p(pkey), _p(ppay), pkey.size, _p(op), _p(ob), k)
    return op, ob


def _aggr(fn, key, val):
    key = np.ascontiguousarray(key, dtype=np.uint32)
    val = np.ascontiguousarray(val, dtype=np.int64)
    n = key.size
    ok = np.empty(n, np.uint32)
    os_ = np.empty(n, np.int64)
    oc = np.empty(n, np.int64)
    g = getattr(lib(), fn)(_p(key), _p(val), n, _p(ok), _p(os_), _p(oc), n)
    return ok[:g], os_[:g], oc[:g]


def hash_aggr(key, val):
    return _aggr("orc_hash_aggr_u32_i64", key, val)


def sort_aggr(key, val):
    return _aggr("orc_sort_aggr_u32_i64", key, val)


Q1_FIELDS = ("key", "sum_qty", "sum_price", "sum_disc", "sum_disc_price", "sum_charge", "count")


def _q1_out(cap):
    return {"key": np.empty(cap, np.int32), "sum_qty": np.empty(cap, np.int64), "sum_price": np.empty(cap, np.int64),
            "sum_disc": np.empty(cap, np.int64), "sum_disc_price": np.empty(cap, np.int64),
            "sum_charge": np.empty(2 * cap, np.int64), "count": np.empty(cap, np.int64)}


def _q1_trim(o, g):
    out = {k: v[:g] for k, v in o.items() if k != "sum_charge"}
    ch = o["sum_charge"][: 2 * g].reshape(g, 2)
    out["sum_charge"] = [int(lo) % (1 << 64) + (int(hi) << 64) for lo, hi in ch]
    return out


def q1_flags(t, cutoff, cap=65536):
    o = _q1_out(cap)
    c = {k: np.ascontiguousarray(t[k]) for k in ("date", "flag1", "flag2", "quantity", "price", "discount", "tax")}
    g = lib().orc_q1_flags(_p(c["date"]), _p(c["flag1"]), _p(c["flag2"]), _p(c["quantity"]), _p(c["price"]),
                           _p(c["discount"]), _p(c["tax"]), c["date"].size, cutoff, cap, _p(o["key"]),
                           _p(o["sum_qty"]), _p(o["sum_price"]), _p(o["sum_disc"]), _p(o["sum_disc_price"]),
                           _p(o["sum_charge"]), _p(o["count"]))
    return _q1_trim(o, g)


def q1_key(t, cutoff):
    n = t["date"].size
    o = _q1_out(max(n, 1))
    c = {k: np.ascontiguousarray(t[k]) for k in ("date", "orderkey", "quantity", "price", "discount", "tax")}
    g = lib().orc_q1_key(_p(c["date"]), _p(c["orderkey"]), _p(c["quantity"]), _p(c["price"]), _p(c["discount"]),
                         _p(c["tax"]), n, cutoff, max(n, 1), _p(o["key"]), _p(o["sum_qty"]), _p(o["sum_price"]),
                         _p(o["sum_disc"]), _p(o["sum_disc_price"]), _p(o["sum_charge"]), _p(o["count"]))
    return _q1_trim(o, g)


# TPC-H Q6 defaults on the fact table's date domain (days from 1992-01-01):
# one year from 1994-01-01, discount 6 +- 1 (hundredths), quantity < 24
Q6_DEFAULT = {"date_lo": 731, "date_hi": 1096, "disc_lo": 5, "disc_hi": 7, "qty_lt": 24}


def q6(t, date_lo=731, date_hi=1096, disc_lo=5, disc_hi=7, qty_lt=24):
    """(revenue, rows kept): revenue = sum(price * discount) over the kept rows, int64 modulo 2^64."""
    c = {k: np.ascontiguousarray(t[k]) for k in ("date", "quantity", "price", "discount")}
    rev = C.c_int64()
    cnt = lib().orc_q6(_p(c["date"]), _p(c["quantity"]), _p(c["price"]), _p(c["discount"]), c["date"].size,
                       date_lo, date_hi, disc_lo, disc_hi, qty_lt, C.byref(rev))
    return rev.value, cnt


def date_pack_str(text: str):
    """Packed date of a "YYYY-MM-DD" literal via strptime, or None when strptime rejects it."""
    out = C.c_int32()
    return out.value if lib().orc_date_pack_str(text.encode(), C.byref(out)) == 0 else None


def date32_pack(days):
    """Packed dates of date32 values (int32 days since 1970-01-01)."""
    days = np.ascontiguousarray(days, dtype=np.int32)
    out = np.empty(days.size, np.int32)
    lib().orc_date32_pack(_p(days), days.size, _p(out))
    return out


def bloom_build(keys, part_bits, shift):
    """Bloom filters of the build keys: (1 << part_bits) x 2048 bytes."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    out = np.empty((1 << part_bits) * 2048, np.uint8)
    lib().orc_bloom_build(_p(keys), keys.size, part_bits, shift, _p(out))
    return out


def bloom_probe(keys, pay, part_bits, shift, filters):
    """(payload, row index) of the probe rows the filters keep, in row order."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    pay = None if pay is None else np.ascontiguousarray(pay, dtype=np.uint32)
    op = np.empty(keys.size, np.uint32)
    orow = np.empty(keys.size, np.uint32)
    k = lib().orc_bloom_probe(_p(keys), _p(pay), keys.size, part_bits, shift, _p(filters), _p(op), _p(orow))
    return op[:k], orow[:k]


def pim_hash(which, key):
    """hash0 (which = 0) or hash1..hash4 of pimdal/hash/hash_func.h."""
    return lib().orc_pim_hash(which, key)


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(t: int) -> None:
    lib().orc_set_threads(t)


# ---------------- host versions of the workload recipes ----------------

SEEDS = {"select": 0x5EED0001, "sort": 0x5EED0002, "join": 0x5EED0003, "groupby": 0x5EED0004, "serve": 0x5EED0005,
         "aggr": 0x5EED0007}
COL = dict(KEY=1, PAYLOAD=2, BUILD_PERM=3, PROBE_FK=4, BUILD_PAY=5, PROBE_PAY=6, ZIPF=7, QTY=12, PRICE=13,
           DISC=14, TAX=15, DATE=16, FLAG1=17, FLAG2=18)


def select_workload(n, seed=SEEDS["select"]):
    return gen_u32(seed, COL["KEY"], n), gen_u32(seed, COL["PAYLOAD"], n)


def sort_workload(n, zipf=False, seed=SEEDS["sort"], distinct=1 << 24):
    if zipf:
        k = fmix32(gen_zipf(seed, COL["ZIPF"], distinct, 1.0, n))
    else:
        k = gen_u32(seed, COL["KEY"], n)
    return k, np.arange(n, dtype=np.uint32)


def join_workload(nr, ns, zipf=False, seed=SEEDS["join"]):
    rk = gen_perm(seed, COL["BUILD_PERM"], nr, nr)
    rp = gen_u32(seed, COL["BUILD_PAY"], nr)
    if zipf:
        ranks = gen_zipf(seed, COL["ZIPF"], nr, 1.0, ns)
        sk = perm_apply(seed, COL["BUILD_PERM"], nr, ranks - np.uint32(1))
    else:
        sk = gen_range(np.uint32, seed, COL["PROBE_FK"], ns, 0, nr - 1)
    sp = gen_u32(seed, COL["PROBE_PAY"], ns)
    return rk, rp, sk, sp


def aggr_workload(n, lo, hi, seed=SEEDS["aggr"]):
    return gen_range(np.uint32, seed, COL["KEY"], n, lo, hi), gen_range(np.int64, seed, COL["PAYLOAD"], n, 1, 10)


def fact_table(n, seed=SEEDS["groupby"], groups=None):
    if groups is None:
        groups = max(1, n // 4)
    return {
        "orderkey": gen_orderkey(seed, n, groups),
        "quantity": gen_range(np.int64, seed, COL["QTY"], n, 1, 50),
        "price": gen_range(np.int64, seed, COL["PRICE"], n, 90000, 10500000),
        "discount": gen_range(np.int64, seed, COL["DISC"], n, 0, 10),
        "tax": gen_range(np.int64, seed, COL["TAX"], n, 0, 8),
        "date": gen_range(np.int32, seed, COL["DATE"], n, 0, 2556),
        "flag1": gen_range(np.uint8, seed, COL["FLAG1"], n, 0, 2),
        "flag2": gen_range(np.uint8, seed, COL["FLAG2"], n, 0, 1),
    }


Q3_COLS = {"customer": ("c_custkey", "c_mktsegment"),
           "orders": ("o_orderkey", "o_custkey", "o_orderdate", "o_shippriority"),
           "lineitem": ("l_orderkey", "l_extendedprice", "l_discount", "l_shipdate")}


def q3(t, segment, date, cap=None):
    """TPC-H Q3 groups in TPC-H order: dict of key, revenue, orderdate,
    shippriority (numpy arrays) and the group count."""
    c = {k: np.ascontiguousarray(t[k]) for cols in Q3_COLS.values() for k in cols}
    nc, no, nl = c["c_custkey"].size, c["o_orderkey"].size, c["l_orderkey"].size
    cap = no if cap is None else cap
    o = {"key": np.empty(max(cap, 1), np.int32), "revenue": np.empty(max(cap, 1), np.int64),
         "orderdate": np.empty(max(cap, 1), np.int32), "shippriority": np.empty(max(cap, 1), np.int32)}
    g = lib().orc_q3(_p(c["c_custkey"]), _p(c["c_mktsegment"]), nc, _p(c["o_orderkey"]), _p(c["o_custkey"]),
                     _p(c["o_orderdate"]), _p(c["o_shippriority"]), no, _p(c["l_orderkey"]), _p(c["l_extendedprice"]),
                     _p(c["l_discount"]), _p(c["l_shipdate"]), nl, segment, date, cap, _p(o["key"]), _p(o["revenue"]),
                     _p(o["orderdate"]), _p(o["shippriority"]))
    if g < 0:
        raise ValueError("o_orderkey values repeat")
    k = min(g, cap)
    return {f: v[:k] for f, v in o.items()}, g


Q4_COLS = {"orders": ("o_orderkey", "o_orderdate", "o_orderpriority"),
           "lineitem": ("l_orderkey", "l_commitdate", "l_receiptdate")}


def q4(t, date_start, date_end):
    """TPC-H Q4: (priority codes with a count, in code order; their order
    counts)."""
    c = {k: np.ascontiguousarray(t[k]) for cols in Q4_COLS.values() for k in cols}
    cnt = np.zeros(256, np.int64)
    g = lib().orc_q4(_p(c["o_orderkey"]), _p(c["o_orderdate"]), _p(c["o_orderpriority"]), c["o_orderkey"].size,
                     _p(c["l_orderkey"]), _p(c["l_commitdate"]), _p(c["l_receiptdate"]), c["l_orderkey"].size,
                     int(date_start), int(date_end), _p(cnt))
    if g < 0:
        raise ValueError("o_orderkey values repeat")
    codes = np.nonzero(cnt)[0].astype(np.uint8)
    return codes, cnt[codes]


Q5_COLS = {"region": ("r_regionkey", "r_name"), "nation": ("n_nationkey", "n_regionkey"),
           "customer": ("c_custkey", "c_nationkey"), "orders": ("o_orderkey", "o_custkey", "o_orderdate"),
           "lineitem": ("l_orderkey", "l_suppkey", "l_extendedprice", "l_discount"),
           "supplier": ("s_suppkey", "s_nationkey")}


def q5(t, region_name, date_start, date_end, cap=1024):
    """TPC-H Q5: (nation keys, revenues), revenue descending then key."""
    c = {k: np.ascontiguousarray(t[k]) for cols in Q5_COLS.values() for k in cols}
    onat, orev = np.empty(max(cap, 1), np.int32), np.empty(max(cap, 1), np.int64)
    n = lambda k: c[k].size
    g = lib().orc_q5(_p(c["r_regionkey"]), _p(c["r_name"]), n("r_regionkey"), _p(c["n_nationkey"]),
                     _p(c["n_regionkey"]), n("n_nationkey"), _p(c["c_custkey"]), _p(c["c_nationkey"]), n("c_custkey"),
                     _p(c["o_orderkey"]), _p(c["o_custkey"]), _p(c["o_orderdate"]), n("o_orderkey"),
                     _p(c["l_orderkey"]), _p(c["l_suppkey"]), _p(c["l_extendedprice"]), _p(c["l_discount"]),
                     n("l_orderkey"), _p(c["s_suppkey"]), _p(c["s_nationkey"]), n("s_suppkey"), int(region_name),
                     int(date_start), int(date_end), cap, _p(onat), _p(orev))
    if g < 0:
        raise ValueError("a key repeats")
    k = min(g, cap)
    return onat[:k], orev[:k]
