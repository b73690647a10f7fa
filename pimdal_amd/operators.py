"""Operator calls over device buffers (the `*_dev` forms of the C ABI).

These mirror PIMDAL's operator set — select, sort, hash join, sort-merge
join, hash and sort aggregation, the fused Q1 filter + group-by — with
device-resident inputs and outputs.  Each call issues HIP kernels through
libpimdal_gpu.so; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call
from .device import Context, DeviceArray, pimdal_type


def _read_count(ctx: Context, d: DeviceArray) -> int:
    ctx.sync()
    return int(d.numpy()[0])


def select(ctx: Context, key: DeviceArray, payload: DeviceArray | None, lo: int, hi: int,
           want_payload: bool = True, want_rowid: bool = True, sync: bool = True):
    """Rows with lo <= key < hi, in row order: (payload, rowid, count).

    reference/db_operators/select.cpp:77-112; pimdal/select/sel.c:59-128.
    """
    n = key.n
    out_p = ctx.empty(n, np.uint32) if want_payload else None
    out_r = ctx.empty(n, np.uint32) if want_rowid else None
    cnt = ctx.empty(1, np.uint64)
    call("pimdal_select_dev", ctx.handle, pimdal_type(key.dtype), key.vp,
         payload.vp if payload is not None else None, n, int(lo), int(hi),
         out_p.vp if out_p is not None else None, out_r.vp if out_r is not None else None, cnt.vp)
    if not sync:
        return out_p, out_r, cnt
    return out_p, out_r, _read_count(ctx, cnt)


def sort_pairs(ctx: Context, keys: DeviceArray, vals: DeviceArray | None, begin_bit: int = -1, end_bit: int = -1,
               out_keys: DeviceArray | None = None, out_vals: DeviceArray | None = None):
    """Stable LSD radix sort of (uint32 key, payload).

    reference/db_operators/sort.cpp:30-70; pimdal/sort/sort_keyval.c:142-179.
    """
    n = keys.n
    ok = out_keys if out_keys is not None else ctx.empty(n, np.uint32)
    vw = 0 if vals is None else vals.dtype.itemsize
    ov = None
    if vals is not None:
        ov = out_vals if out_vals is not None else ctx.empty(n, vals.dtype)
    call("pimdal_sort_pairs_dev", ctx.handle, keys.vp, vals.vp if vals is not None else None, vw, n, ok.vp,
         ov.vp if ov is not None else None, begin_bit, end_bit)
    return ok, ov


def _join(name: str, ctx: Context, bk, bp, pk, pp, capacity: int | None = None):
    cap = pk.n if capacity is None else int(capacity)
    op = ctx.empty(max(cap, 1), np.uint32)
    ob = ctx.empty(max(cap, 1), np.uint32)
    cnt = C.c_uint64()
    rc = getattr(_lib.load(), name)(ctx.handle, bk.vp, bp.vp if bp is not None else None, bk.n, pk.vp,
                                    pp.vp if pp is not None else None, pk.n, op.vp, ob.vp, cap, C.byref(cnt))
    if rc == -5 and capacity is None:  # more pairs than probe rows: size exactly and rerun
        op.free(); ob.free()
        return _join(name, ctx, bk, bp, pk, pp, cnt.value)
    _lib.check(rc)
    return op, ob, cnt.value


def hash_join(ctx: Context, build_key, build_payload, probe_key, probe_payload, capacity: int | None = None):
    """(probe payload, build payload) of every matching pair, unordered.

    reference/db_operators/hash_join.cpp:159-219; pimdal/join/hash_join.c:40-430.
    """
    return _join("pimdal_hash_join_dev", ctx, build_key, build_payload, probe_key, probe_payload, capacity)


def sort_merge_join(ctx: Context, build_key, build_payload, probe_key, probe_payload, capacity: int | None = None):
    """Every matching pair in (key, probe row, build row) order.

    reference/db_operators/sort_merge_join.cpp:48-156; pimdal/join/sort_merge.c:54-152.
    """
    return _join("pimdal_sort_merge_join_dev", ctx, build_key, build_payload, probe_key, probe_payload, capacity)


def _aggr(name: str, ctx: Context, key, val, capacity: int | None = None):
    cap = max(key.n if capacity is None else int(capacity), 1)
    ok = ctx.empty(cap, np.uint32)
    os_ = ctx.empty(cap, np.int64)
    oc = ctx.empty(cap, np.int64)
    g = C.c_uint64()
    call(name, ctx.handle, key.vp, val.vp, key.n, ok.vp, os_.vp, oc.vp, cap, C.byref(g))
    return ok, os_, oc, g.value


def hash_aggr(ctx: Context, key, val, capacity: int | None = None):
    """(key, SUM(val), COUNT) per group, unordered.

    reference/db_operators/aggregate.cpp:40-151; pimdal/aggregate/aggregate_hash.c:43-193.
    """
    return _aggr("pimdal_hash_aggr_dev", ctx, key, val, capacity)


def sort_aggr(ctx: Context, key, val, capacity: int | None = None):
    """(key, SUM(val), COUNT) per group in ascending key order.

    pimdal/aggregate/aggregate.c:31-158.
    """
    return _aggr("pimdal_sort_aggr_dev", ctx, key, val, capacity)


def q6(ctx: Context, t: dict, date_lo: int = 731, date_hi: int = 1096, disc_lo: int = 5, disc_hi: int = 7,
       qty_lt: int = 24):
    """TPC-H Q6 shape: (revenue, rows kept), revenue = sum(price * discount)
    over rows with date_lo <= date < date_hi, quantity < qty_lt and
    disc_lo <= discount <= disc_hi (int64 modulo 2^64).
    pimdal/main/tpc_h/query6/dpu/kernel_q6.c:91-141, host_q6.cpp:46-49.
    """
    rev = C.c_int64()
    cnt = C.c_uint64()
    call("pimdal_q6_dev", ctx.handle, t["date"].vp, t["quantity"].vp, t["price"].vp, t["discount"].vp,
         t["date"].n, int(date_lo), int(date_hi), int(disc_lo), int(disc_hi), int(qty_lt), C.byref(rev),
         C.byref(cnt))
    return rev.value, cnt.value


def bloom_build(ctx: Context, keys: DeviceArray, part_bits: int, shift: int) -> DeviceArray:
    """Bloom filters of the build keys, (1 << part_bits) x 2048 bytes (as uint32 words).
    pimdal/join/hash_join.c:117-152 (filter_kernel)."""
    filters = ctx.empty(512 << part_bits, np.uint32)
    call("pimdal_bloom_build_dev", ctx.handle, keys.vp, keys.n, int(part_bits), int(shift), filters.vp)
    return filters


def bloom_probe(ctx: Context, keys: DeviceArray, payload: DeviceArray | None, part_bits: int, shift: int,
                filters: DeviceArray):
    """Probe rows the filters keep, in row order: (payload, rowid, count).
    pimdal/join/hash_join.c:237-341 (match_kernel)."""
    n = keys.n
    out_p, out_r, cnt = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32), ctx.empty(1, np.uint64)
    call("pimdal_bloom_probe_dev", ctx.handle, keys.vp, payload.vp if payload is not None else None, n,
         int(part_bits), int(shift), filters.vp, out_p.vp, out_r.vp, cnt.vp)
    return out_p, out_r, _read_count(ctx, cnt)


def bloom_semi_join(ctx: Context, build_keys: DeviceArray, probe_keys: DeviceArray, probe_payload: DeviceArray | None,
                    part_bits: int, shift: int):
    """Build the filters and probe them in one call: (payload, rowid, count)."""
    n = probe_keys.n
    out_p, out_r, cnt = ctx.empty(n, np.uint32), ctx.empty(n, np.uint32), ctx.empty(1, np.uint64)
    call("pimdal_bloom_semi_join_dev", ctx.handle, build_keys.vp, build_keys.n, probe_keys.vp,
         probe_payload.vp if probe_payload is not None else None, n, int(part_bits), int(shift), out_p.vp,
         out_r.vp, cnt.vp)
    return out_p, out_r, _read_count(ctx, cnt)


def date_pack(year: int, month: int, day: int) -> int:
    """Packed date day + (month << 5) + ((year - 1900) << 9):
    pimdal/main/tpc_h/reader/read_table.cpp:8-16 (date_to_int)."""
    out = C.c_int32()
    call("pimdal_date_pack", int(year), int(month), int(day), C.byref(out))
    return out.value


def date_pack_str(text: str) -> int:
    """Packed date of a "YYYY-MM-DD" literal (read_table.cpp:8-16)."""
    out = C.c_int32()
    call("pimdal_date_pack_str", text.encode(), C.byref(out))
    return out.value


def date32_pack(ctx: Context, days: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
    """Arrow date32 values (int32 days since 1970-01-01) to packed dates on
    the device; `out` may be `days` itself (in place)."""
    if days.dtype != np.int32:
        raise TypeError("date32 values are int32")
    if out is None:
        out = ctx.empty(days.n, np.int32)
    call("pimdal_date32_pack_dev", ctx.handle, days.vp, days.n, out.vp)
    return out


class _Q1Result(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in ("key", "sum_qty", "sum_price", "sum_disc", "sum_disc_price",
                                          "sum_charge", "count")]


Q1_FIELDS = ("key", "sum_qty", "sum_price", "sum_disc", "sum_disc_price", "sum_charge", "count")


def filter_groupby_q1(ctx: Context, t: dict, cutoff: int, by_key: bool = False, capacity: int | None = None,
                      to_host: bool = True):
    """Fused date <= cutoff filter and group-by with the TPC-H Q1 measures.

    Groups by (flag1, flag2) or, with by_key, by t["orderkey"].
    pimdal/main/tpc_h/query1/dpu/kernel_q1_1.c:59-228, kernel_q1_2.c:37-46.
    """
    n = t["date"].n
    cap = max(1, (64 if not by_key else n) if capacity is None else int(capacity))
    out = {f: ctx.empty(cap * (2 if f == "sum_charge" else 1), np.int32 if f == "key" else np.int64)
           for f in Q1_FIELDS}
    res = _Q1Result(*[out[f].ptr for f in Q1_FIELDS])
    g = C.c_uint64()
    call("pimdal_filter_groupby_q1_dev", ctx.handle, t["date"].vp, t["flag1"].vp, t["flag2"].vp,
         t["orderkey"].vp if by_key else None, t["quantity"].vp, t["price"].vp, t["discount"].vp, t["tax"].vp,
         n, int(cutoff), C.byref(res), cap, C.byref(g))
    k = g.value
    if not to_host:
        return out, k
    h = {f: out[f].numpy(2 * k if f == "sum_charge" else k) for f in Q1_FIELDS}
    ch = h["sum_charge"].reshape(k, 2)
    h["sum_charge"] = [int(lo) % (1 << 64) + (int(hi) << 64) for lo, hi in ch]
    return h


def _q1_outputs(ctx: Context, cap: int):
    out = {f: ctx.empty(cap * (2 if f == "sum_charge" else 1), np.int32 if f == "key" else np.int64)
           for f in Q1_FIELDS}
    return out, _Q1Result(*[out[f].ptr for f in Q1_FIELDS])


def filter_groupby_q1_async(ctx: Context, t: dict, cutoff: int, by_key: bool = False, capacity: int | None = None):
    """The asynchronous Q1 form (pimdal_filter_groupby_q1_dev_async): enqueues
    the work and returns (outputs, d_groups, d_status) as device arrays;
    nothing is read back, so it can run inside a Plan."""
    n = t["date"].n
    cap = max(1, (64 if not by_key else n) if capacity is None else int(capacity))
    out, res = _q1_outputs(ctx, cap)
    d_ng, d_st = ctx.empty(1, np.uint64), ctx.empty(1, np.int32)
    out["_res"] = res   # the descriptor is read at the call only; kept with the outputs for clarity
    call("pimdal_filter_groupby_q1_dev_async", ctx.handle, t["date"].vp, t["flag1"].vp, t["flag2"].vp,
         t["orderkey"].vp if by_key else None, t["quantity"].vp, t["price"].vp, t["discount"].vp, t["tax"].vp,
         n, int(cutoff), C.byref(res), cap, d_ng.vp, d_st.vp)
    return out, d_ng, d_st


def q1_to_host(out: dict, d_ng: DeviceArray, d_st: DeviceArray):
    """(status, host result) of an asynchronous Q1 call once its stream has run
    (status 0, or the negative PIMDAL_E_* code the call left)."""
    st = int(d_st.numpy()[0])
    k = int(d_ng.numpy()[0])
    cap = out["key"].n
    kk = min(k, cap)
    h = {f: out[f].numpy(2 * kk if f == "sum_charge" else kk) for f in Q1_FIELDS}
    ch = h["sum_charge"].reshape(kk, 2)
    h["sum_charge"] = [int(lo) % (1 << 64) + (int(hi) << 64) for lo, hi in ch]
    h["groups"] = k
    return st, h


def q6_async(ctx: Context, t: dict, date_lo: int = 731, date_hi: int = 1096, disc_lo: int = 5, disc_hi: int = 7,
             qty_lt: int = 24):
    """pimdal_q6_dev_async: (d_revenue int64[1], d_rows uint64[1]) device words."""
    rev, rows = ctx.empty(1, np.int64), ctx.empty(1, np.uint64)
    call("pimdal_q6_dev_async", ctx.handle, t["date"].vp, t["quantity"].vp, t["price"].vp, t["discount"].vp,
         t["date"].n, int(date_lo), int(date_hi), int(disc_lo), int(disc_hi), int(qty_lt), rev.vp, rows.vp)
    return rev, rows


class Plan:
    """A captured sequence of asynchronous device calls (pimdal_plan_*):

        with Plan(ctx) as plan:       # capture
            ...asynchronous calls...
        plan.launch(); ctx.sync()     # replay

    The calls inside the block are recorded, not run.  The plan keeps the
    scratch memory of the captured calls until destroy()."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.handle = None

    def __enter__(self):
        h = C.c_void_p()
        call("pimdal_plan_begin", self.ctx.handle, C.byref(h))
        self.handle = h
        return self

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None:
            try:
                call("pimdal_plan_end", self.handle)
            except Exception:
                self.destroy()
                raise
        else:
            self.destroy()
        return False

    def launch(self) -> None:
        call("pimdal_plan_launch", self.handle)

    def destroy(self) -> None:
        if self.handle:
            call("pimdal_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                _lib.load().pimdal_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class _Q3Tables(C.Structure):
    _fields_ = [("c_custkey", C.c_void_p), ("c_mktsegment", C.c_void_p), ("n_customer", C.c_uint64),
                ("o_orderkey", C.c_void_p), ("o_custkey", C.c_void_p), ("o_orderdate", C.c_void_p),
                ("o_shippriority", C.c_void_p), ("n_orders", C.c_uint64), ("l_orderkey", C.c_void_p),
                ("l_extendedprice", C.c_void_p), ("l_discount", C.c_void_p), ("l_shipdate", C.c_void_p),
                ("n_lineitem", C.c_uint64)]


class _Q3Result(C.Structure):
    _fields_ = [("orderkey", C.c_void_p), ("revenue", C.c_void_p), ("orderdate", C.c_void_p),
                ("shippriority", C.c_void_p)]


def tpch_q3(ctx: Context, t: dict, segment: int, date: int, capacity: int = 10, to_host: bool = True):
    """TPC-H Q3 shape over device columns `t` (customer, orders, lineitem):
    the first `capacity` groups in TPC-H order and the group count.
    pimdal/main/tpc_h/query3/host/host_q3.cpp:471-519."""
    tab = _Q3Tables(t["c_custkey"].ptr, t["c_mktsegment"].ptr, t["c_custkey"].n, t["o_orderkey"].ptr,
                    t["o_custkey"].ptr, t["o_orderdate"].ptr, t["o_shippriority"].ptr, t["o_orderkey"].n,
                    t["l_orderkey"].ptr, t["l_extendedprice"].ptr, t["l_discount"].ptr, t["l_shipdate"].ptr,
                    t["l_orderkey"].n)
    cap = max(1, int(capacity))
    out = {"key": ctx.empty(cap, np.int32), "revenue": ctx.empty(cap, np.int64),
           "orderdate": ctx.empty(cap, np.int32), "shippriority": ctx.empty(cap, np.int32)}
    res = _Q3Result(out["key"].ptr, out["revenue"].ptr, out["orderdate"].ptr, out["shippriority"].ptr)
    g = C.c_uint64()
    call("pimdal_tpch_q3_dev", ctx.handle, C.byref(tab), int(segment), int(date), C.byref(res), cap, C.byref(g))
    if not to_host:
        return out, g.value
    k = min(g.value, cap)
    return {f: a.numpy(k) for f, a in out.items()}, g.value


class _Q4Tables(C.Structure):
    _fields_ = [("o_orderkey", C.c_void_p), ("o_orderdate", C.c_void_p), ("o_orderpriority", C.c_void_p),
                ("n_orders", C.c_uint64), ("l_orderkey", C.c_void_p), ("l_commitdate", C.c_void_p),
                ("l_receiptdate", C.c_void_p), ("n_lineitem", C.c_uint64)]


def tpch_q4(ctx: Context, t: dict, date_start: int, date_end: int, capacity: int = 256):
    """TPC-H Q4 shape over device columns `t` (orders, lineitem): priority
    codes with at least one order, in code order, their order counts, and
    the group count.  pimdal/main/tpc_h/query4/host/host_q4.cpp:303-330."""
    tab = _Q4Tables(t["o_orderkey"].ptr, t["o_orderdate"].ptr, t["o_orderpriority"].ptr, t["o_orderkey"].n,
                    t["l_orderkey"].ptr, t["l_commitdate"].ptr, t["l_receiptdate"].ptr, t["l_orderkey"].n)
    cap = max(1, int(capacity))
    prio, cnt = ctx.empty(cap, np.uint8), ctx.empty(cap, np.int64)
    g = C.c_uint64()
    call("pimdal_tpch_q4_dev", ctx.handle, C.byref(tab), int(date_start), int(date_end), prio.ptr, cnt.ptr, cap,
         C.byref(g))
    k = min(g.value, cap)
    return prio.numpy(k), cnt.numpy(k), g.value


class _Q5Tables(C.Structure):
    _fields_ = [("r_regionkey", C.c_void_p), ("r_name", C.c_void_p), ("n_region", C.c_uint64),
                ("n_nationkey", C.c_void_p), ("n_regionkey", C.c_void_p), ("n_nation", C.c_uint64),
                ("c_custkey", C.c_void_p), ("c_nationkey", C.c_void_p), ("n_customer", C.c_uint64),
                ("o_orderkey", C.c_void_p), ("o_custkey", C.c_void_p), ("o_orderdate", C.c_void_p),
                ("n_orders", C.c_uint64), ("l_orderkey", C.c_void_p), ("l_suppkey", C.c_void_p),
                ("l_extendedprice", C.c_void_p), ("l_discount", C.c_void_p), ("n_lineitem", C.c_uint64),
                ("s_suppkey", C.c_void_p), ("s_nationkey", C.c_void_p), ("n_supplier", C.c_uint64)]


def tpch_q5(ctx: Context, t: dict, region_name: int, date_start: int, date_end: int, capacity: int = 1024):
    """TPC-H Q5 shape over device columns `t` (region, nation, customer,
    orders, lineitem, supplier): nation keys and revenues, revenue
    descending, and the nation count.
    pimdal/main/tpc_h/query5/host/host_q5.cpp:566-640."""
    tab = _Q5Tables(t["r_regionkey"].ptr, t["r_name"].ptr, t["r_regionkey"].n, t["n_nationkey"].ptr,
                    t["n_regionkey"].ptr, t["n_nationkey"].n, t["c_custkey"].ptr, t["c_nationkey"].ptr,
                    t["c_custkey"].n, t["o_orderkey"].ptr, t["o_custkey"].ptr, t["o_orderdate"].ptr,
                    t["o_orderkey"].n, t["l_orderkey"].ptr, t["l_suppkey"].ptr, t["l_extendedprice"].ptr,
                    t["l_discount"].ptr, t["l_orderkey"].n, t["s_suppkey"].ptr, t["s_nationkey"].ptr,
                    t["s_suppkey"].n)
    cap = max(1, int(capacity))
    nat, rev = ctx.empty(cap, np.int32), ctx.empty(cap, np.int64)
    g = C.c_uint64()
    call("pimdal_tpch_q5_dev", ctx.handle, C.byref(tab), int(region_name), int(date_start), int(date_end), nat.ptr,
         rev.ptr, cap, C.byref(g))
    k = min(g.value, cap)
    return nat.numpy(k), rev.numpy(k), g.value
