#!/usr/bin/env python3
"""Benchmark of the DB-operator hot path on one MI355X.

Headline workload: stable LSD radix sort of 2^28 (uint32 key, uint32 payload)
pairs with keys uniform on [0, 2^32).  One step = one sort of the whole
batch, inputs already resident in HBM.  The JSON line also carries the other
operators' rows/s at their configured sizes, the dominant kernel's roofline
(hipEvent time on the stream it runs on), and the CPU restatement timed on
a bounded sample on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--only sort|zipf|sortselect|join|bloom|groupby|aggr|q3|ingest|serve]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12
METRIC = "input rows/s & %HBM-peak per operator; served queries/s + p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(ctx, fn, steps, warmup):
    """Seconds per step over `steps` uninstrumented steps, timed between two
    device synchronisations; then the same number of steps again with
    per-kernel hipEvent timing on, for kernel_breakdown and the roofline's
    kernel durations (events around every launch would add host time to
    the timed steps)."""
    for _ in range(warmup):
        fn()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    ctx.sync()
    t1 = time.perf_counter()
    ctx.reset_stats()
    ctx.set_profiling(True)
    for _ in range(steps):
        fn()
    ctx.sync()
    ctx.set_profiling(False)
    return (t1 - t0) / steps


def kernel_avg_ms(ctx, name):
    n, ms = ctx.kernel_stat(name)
    return (ms / n if n else 0.0), n


def load_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    from the summary committed under profiles/ (None when absent)."""
    p = ROOT / "profiles" / f"{name}_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("bytes_per_launch")
    except Exception:
        return None


# ---------------- operators ----------------

def bench_sort(ctx, steps, warmup, n=1 << 28, zipf=False):
    from pimdal_amd import operators as ops, datagen
    keys, vals = datagen.sort_workload(ctx, n, zipf=zipf)
    ok = ctx.empty(n, "uint32")
    ov = ctx.empty(n, "uint32")
    step = lambda: ops.sort_pairs(ctx, keys, vals, 0, 32, out_keys=ok, out_vals=ov)
    t = timed_steps(ctx, step, steps, warmup)
    pass_ms, launches = kernel_avg_ms(ctx, "radix_onesweep")
    hist_ms, _ = kernel_avg_ms(ctx, "radix_hist")
    count_ms, count_launches = kernel_avg_ms(ctx, "radix_count")
    tscan_ms, _ = kernel_avg_ms(ctx, "radix_tscan")
    bucket_ms, bucket_launches = kernel_avg_ms(ctx, "radix_bucket")
    for a in (keys, vals, ok, ov):
        a.free()
    # trips through HBM per sort: one per onesweep pass, one for the bucket
    # pass (which sorts the two low digits of every bucket in LDS)
    per_step = lambda launches: launches / steps if launches else 0
    trips = round(per_step(launches) + per_step(bucket_launches))
    alg = (4 + 16 * trips) * n
    return {
        "n": n, "s_per_step": t, "rows_per_s": n / t, "hbm_trips": trips,
        "alg_bytes": alg, "alg_GBps": alg / t / 1e9, "hbm_frac": alg / t / HBM_PEAK,
        "keys_payload_GBps": 8 * n / t / 1e9,
        "pass_ms": pass_ms, "pass_launches": launches, "hist_ms": hist_ms,
        "bucket_ms": bucket_ms, "bucket_launches": bucket_launches,
        "count_ms": count_ms, "count_launches": count_launches, "tscan_ms": tscan_ms,
        "pass_alg_bytes": 16 * n,
    }


def bench_select(ctx, steps, warmup, n=1 << 27):
    from pimdal_amd import operators as ops, datagen
    key, pay = datagen.select_workload(ctx, n)
    out = {}
    res = {}

    def step():
        res["r"] = ops.select(ctx, key, pay, 0, datagen.SELECT_THRESHOLD, sync=False)

    t = timed_steps(ctx, step, steps, warmup)
    cnt = int(res["r"][2].numpy()[0])
    ms, _ = kernel_avg_ms(ctx, "select")
    alg = 8 * n + 8 * cnt
    out.update(n=n, selected=cnt, s_per_step=t, rows_per_s=n / t, alg_bytes=alg,
               kernel_ms=ms, alg_GBps=alg / (ms / 1e3) / 1e9 if ms else None,
               hbm_frac=alg / (ms / 1e3) / HBM_PEAK if ms else None)
    key.free()
    pay.free()
    return out


KERNELS = ("bloom_build", "bloom_flags", "bloom_group", "bloom_or", "date32_pack", "part_hist", "part_count", "part_tscan", "radix_count", "radix_tscan", "part_onesweep", "part_bounds", "scan", "hj_chunks", "hj_join", "hj_join_count",
           "hj_join_write", "hj_compact", "hg_build", "hg_probe", "hg_probe_count", "hg_probe_write", "radix_hist",
           "radix_scan", "radix_onesweep", "smj_dupcheck", "smj_ranges", "smj_merge", "smj_expand", "q1_domain",
           "q1_lut", "q1_count", "q1_dense", "q1_reduce", "q1_final", "gb_direct", "gb_direct_emit", "gb_chunks", "gb_hist",
           "gb_scatter", "gb_lds", "gb_global", "gb_emit_global", "q1_unpack", "agg_heads", "agg_run_reduce",
           "select", "select_count", "select_write", "q1_pack", "q6", "q6_final", "smj_pack", "fill_u32", "fill_u64",
           "q3_cust", "q3_osel", "q3_orders", "q3_lineitem", "q3_emit", "q3_order", "q3_topk", "q4_osel", "q4_orders",
           "q4_bits", "q4_lineitem", "q4_count", "q4_emit", "q5_nations", "q5_cust", "q5_osel", "q5_orders",
           "q5_supp", "q5_lineitem", "q5_emit", "as_table", "as_fold",
           "as_emit", "as_gather", "as_runs", "as_fix", "pw_count", "pw_tscan", "pw_last", "pw_scatter",
           "ad_minmax", "ad_count", "ad_tscan", "ad_scatter", "ad_sum", "ad_direct", "radix_bstart", "radix_chunks",
           "radix_bucket", "radix_bucket_retry", "q1_status")


def kernel_breakdown(ctx, steps):
    out = {}
    for k in KERNELS:
        n, ms = ctx.kernel_stat(k)
        if n:
            out[k] = {"launches_per_step": n / steps, "ms_per_step": ms / steps}
    return out


def bench_join(ctx, steps, warmup, which, zipf, nr=1 << 24, ns=1 << 27):
    from pimdal_amd import operators as ops, datagen
    rk, rp, sk, sp = datagen.join_workload(ctx, nr, ns, zipf=zipf)
    op = ctx.empty(ns, "uint32")
    ob = ctx.empty(ns, "uint32")
    fn = ops.hash_join if which == "hash" else ops.sort_merge_join
    import ctypes as C
    from pimdal_amd import _lib
    name = "pimdal_hash_join_dev" if which == "hash" else "pimdal_sort_merge_join_dev"
    cnt = C.c_uint64()

    def step():
        _lib.check(getattr(_lib.load(), name)(ctx.handle, rk.vp, rp.vp, nr, sk.vp, sp.vp, ns, op.vp, ob.vp, ns,
                                              C.byref(cnt)))

    t = timed_steps(ctx, step, steps, warmup)
    alg = (24 * nr + 32 * ns) if which == "hash" else (60 * nr + 68 * ns)
    res = {"build_rows": nr, "probe_rows": ns, "pairs": cnt.value, "s_per_step": t, "rows_per_s": (nr + ns) / t,
           "alg_bytes": alg, "alg_GBps": alg / t / 1e9, "hbm_frac": alg / t / HBM_PEAK,
           "kernels": kernel_breakdown(ctx, steps)}
    for a in (rk, rp, sk, sp, op, ob):
        a.free()
    return res


def bench_q1(ctx, steps, warmup, by_key, n=60_000_000):
    from pimdal_amd import operators as ops, datagen
    t = datagen.fact_table(ctx, n, groups=15_000_000)
    out = {}

    def step():
        out["r"] = ops.filter_groupby_q1(ctx, t, 2505, by_key=by_key, to_host=False)

    tm = timed_steps(ctx, step, steps, warmup)
    groups = out["r"][1]
    nsel = None
    if by_key:
        alg = 42 * n + 88 * int(0.98 * n) + 52 * groups
    else:
        alg = 38 * n
    res = {"rows": n, "groups": groups, "s_per_step": tm, "rows_per_s": n / tm, "alg_bytes": alg,
           "alg_GBps": alg / tm / 1e9, "hbm_frac": alg / tm / HBM_PEAK, "kernels": kernel_breakdown(ctx, steps)}
    for a in t.values():
        a.free()
    return res


def bench_q1_plan(ctx, steps, warmup, by_key, n=60_000_000):
    """The same pipeline captured once as a plan (one hipGraph) and replayed:
    a step is one graph launch.  By key, the partition fan-out and table size
    are autotuned at this size first (pimdal_autotune "q1")."""
    from pimdal_amd import operators as ops, datagen
    t = datagen.fact_table(ctx, n, groups=15_000_000)
    tuned = {}
    if by_key:
        ctx.autotune("q1", n)
        tuned = {k: ctx.get_tune(k) for k in ("gb_rows_per_part", "gb_q1_slots")}
    with ops.Plan(ctx) as plan:
        out, ng, st = ops.filter_groupby_q1_async(ctx, t, 2505, by_key=by_key)
    for k in tuned:
        ctx.clear_tune(k)   # baked into the captured graph; later operators take their defaults
    for _ in range(warmup):
        plan.launch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.launch()
    ctx.sync()
    tm = (time.perf_counter() - t0) / steps
    status = int(st.numpy()[0])
    groups = int(ng.numpy()[0])
    plan.destroy()
    alg = 42 * n + 88 * int(0.98 * n) + 52 * groups if by_key else 38 * n
    res = {"rows": n, "groups": groups, "status": status, "s_per_step": tm, "rows_per_s": n / tm, "alg_bytes": alg,
           "alg_GBps": alg / tm / 1e9, "hbm_frac": alg / tm / HBM_PEAK, "graph_launches_per_step": 1,
           "autotuned": tuned}
    for a in list(t.values()) + [a for k, a in out.items() if k != "_res"] + [ng, st]:
        a.free()
    return res


def bench_q3(ctx, steps, warmup, n_lineitem=60_000_000):
    """TPC-H Q3 shape at SF10's sizes (6e7 lineitems, 1.5e7 orders, 1.5e6
    customers; host-generated synthetic tables copied to HBM before timing):
    customer set, orders table, lineitem probe with the revenue summed in the
    table, all groups in TPC-H order.  Algorithmic bytes: each input column
    read once, the two order columns of the filter twice (count, then build)."""
    import numpy as np
    from pimdal_amd import operators as ops, datagen
    t = datagen.q3_tables(n_lineitem)
    d = {k: ctx.to_device(v) for k, v in t.items()}
    date = int(datagen.pack_days(np.array([9204]))[0])
    seg = datagen.Q3_SEGMENTS.index("BUILDING")
    nc, no, nl = t["c_custkey"].size, t["o_orderkey"].size, n_lineitem
    out = {}

    def step():
        out["r"] = ops.tpch_q3(ctx, d, seg, date, capacity=10, to_host=False)

    tm = timed_steps(ctx, step, steps, warmup)
    groups = out["r"][1]
    alg = 5 * nc + 20 * no + 24 * nl
    res = {"lineitem": nl, "orders": no, "customer": nc, "groups": groups, "s_per_step": tm,
           "rows_per_s": (nc + no + nl) / tm, "alg_bytes": alg, "alg_GBps": alg / tm / 1e9, "hbm_frac": alg / tm / HBM_PEAK,
           "kernels": kernel_breakdown(ctx, steps)}
    for a in list(d.values()) + list(out["r"][0].values()):
        a.free()
    return res


def bench_q4(ctx, steps, warmup, n_lineitem=60_000_000):
    """TPC-H Q4 shape at SF10's sizes (6e7 lineitems, 1.5e7 orders;
    host-generated synthetic tables copied to HBM before timing): orders of
    1993-07-01 .. 1993-09-30 in a table, one lineitem pass marking the
    orders with a late lineitem, counts per priority.  Algorithmic bytes:
    each input column read once (orders 9 B, lineitems 12 B per row)."""
    import numpy as np
    from pimdal_amd import operators as ops, datagen
    t = datagen.q4_tables(n_lineitem)
    d = {k: ctx.to_device(v) for k, v in t.items()}
    lo, hi = (int(x) for x in datagen.pack_days(np.array([8582, 8674])))
    no, nl = t["o_orderkey"].size, n_lineitem
    out = {}

    def step():
        out["r"] = ops.tpch_q4(ctx, d, lo, hi)

    tm = timed_steps(ctx, step, steps, warmup)
    alg = 9 * no + 12 * nl
    res = {"lineitem": nl, "orders": no, "groups": out["r"][2], "s_per_step": tm, "rows_per_s": (no + nl) / tm,
           "alg_bytes": alg, "alg_GBps": alg / tm / 1e9, "hbm_frac": alg / tm / HBM_PEAK,
           "kernels": kernel_breakdown(ctx, steps)}
    for a in d.values():
        a.free()
    return res


def bench_q5(ctx, steps, warmup, n_lineitem=60_000_000):
    """TPC-H Q5 shape at SF10's sizes (6e7 lineitems, 1.5e7 orders, 1.5e6
    customers, 1e5 suppliers; host-generated synthetic tables copied to HBM
    before timing): region ASIA, orders of 1994.  Algorithmic bytes: each
    input column read once (customers 8 B, orders 12 B, lineitems 24 B,
    suppliers 8 B per row)."""
    import numpy as np
    from pimdal_amd import operators as ops, datagen
    t = datagen.q5_tables(n_lineitem)
    d = {k: ctx.to_device(v) for k, v in t.items()}
    lo, hi = (int(x) for x in datagen.pack_days(np.array([8766, 9131])))
    asia = datagen.Q5_REGIONS.index("ASIA")
    nc, no, ns, nl = t["c_custkey"].size, t["o_orderkey"].size, t["s_suppkey"].size, n_lineitem
    out = {}

    def step():
        out["r"] = ops.tpch_q5(ctx, d, asia, lo, hi)

    tm = timed_steps(ctx, step, steps, warmup)
    alg = 8 * nc + 12 * no + 24 * nl + 8 * ns
    res = {"lineitem": nl, "orders": no, "customer": nc, "supplier": ns, "groups": out["r"][2], "s_per_step": tm,
           "rows_per_s": (nc + no + nl + ns) / tm, "alg_bytes": alg, "alg_GBps": alg / tm / 1e9,
           "hbm_frac": alg / tm / HBM_PEAK, "kernels": kernel_breakdown(ctx, steps)}
    for a in d.values():
        a.free()
    return res


def bench_date32(ctx, steps, warmup, n=60_000_000):
    """Arrow date32 column (days over 1992-1998) to the reference's packed dates, 6e7 values."""
    from pimdal_amd import operators as ops
    days = ctx.empty(n, "int32")
    from pimdal_amd._lib import call
    call("pimdal_gen_range", ctx.handle, 2, 0xDA7E, 1, n, 8035, 8035 + 2556, days.vp)
    out = ctx.empty(n, "int32")

    def step():
        ops.date32_pack(ctx, days, out)

    tm = timed_steps(ctx, step, steps, warmup)
    alg = 8 * n   # 4 B read + 4 B written per value
    res = {"rows": n, "s_per_step": tm, "rows_per_s": n / tm, "alg_bytes": alg, "alg_GBps": alg / tm / 1e9,
           "hbm_frac": alg / tm / HBM_PEAK, "kernels": kernel_breakdown(ctx, steps)}
    days.free()
    out.free()
    return res


def bench_bloom(ctx, steps, warmup, nb, part_bits, npr=1 << 27):
    """Bloom-filter semi-join: nb build keys into 2^part_bits filters, then a
    probe of npr keys with payloads (a quarter of them build keys)."""
    from pimdal_amd import operators as ops
    from pimdal_amd._lib import call
    import ctypes as C
    bk = ctx.empty(nb, "uint32")
    call("pimdal_gen_u32", ctx.handle, 0xB10, 1, nb, bk.vp)
    pk, pp = ctx.empty(npr, "uint32"), ctx.empty(npr, "uint32")
    call("pimdal_gen_u32", ctx.handle, 0xB10, 2, npr, pk.vp)
    call("pimdal_gen_u32", ctx.handle, 0xB10, 3, npr, pp.vp)
    for q in range(0, npr, 4 * nb):       # every 4th block of nb probe keys repeats the build keys
        m = min(nb, npr - q)
        call("pimdal_dev_copy", ctx.handle, C.c_void_p(pk.ptr + 4 * q), bk.vp, 4 * m)
    op, orow, cnt = ctx.empty(npr, "uint32"), ctx.empty(npr, "uint32"), ctx.empty(1, "uint64")

    def step():
        call("pimdal_bloom_semi_join_dev", ctx.handle, bk.vp, nb, pk.vp, pp.vp, npr, part_bits, 32, op.vp,
             orow.vp, cnt.vp)

    tm = timed_steps(ctx, step, steps, warmup)
    kept = int(cnt.numpy()[0])
    alg = 4 * nb + 8 * npr + 8 * kept    # build keys; probe key + payload; kept payload + row id
    res = {"build_rows": nb, "probe_rows": npr, "filters": 1 << part_bits, "kept": kept, "s_per_step": tm,
           "rows_per_s": (nb + npr) / tm, "alg_bytes": alg, "alg_GBps": alg / tm / 1e9, "hbm_frac": alg / tm / HBM_PEAK,
           "kernels": kernel_breakdown(ctx, steps)}
    for a in (bk, pk, pp, op, orow, cnt):
        a.free()
    return res


def bench_q6(ctx, steps, warmup, n=60_000_000):
    """TPC-H Q6 shape over the 6e7-row fact table: one year of dates, discount 5..7, quantity < 24."""
    from pimdal_amd import operators as ops, datagen
    t = datagen.fact_table(ctx, n, groups=15_000_000)
    out = {}

    def step():
        out["r"] = ops.q6(ctx, t)

    tm = timed_steps(ctx, step, steps, warmup)
    alg = 28 * n   # date 4 + quantity 8 + price 8 + discount 8 bytes per row
    res = {"rows": n, "rows_kept": out["r"][1], "s_per_step": tm, "rows_per_s": n / tm, "alg_bytes": alg,
           "alg_GBps": alg / tm / 1e9, "hbm_frac": alg / tm / HBM_PEAK, "kernels": kernel_breakdown(ctx, steps)}
    for a in t.values():
        a.free()
    return res


AGGR_SHAPES = {"k50": (1, 50), "k4000": (1, 4000), "k25": (1, 25), "k2^24": (0, (1 << 24) - 1)}


def bench_aggr(ctx, steps, warmup, which, shape, n=1 << 27):
    """Hash or sort aggregation (SUM and COUNT per uint32 key) of n rows with
    keys uniform over the shape's range and int64 values on [1, 10].
    Algorithmic bytes: key and value read once, 20 bytes out per group, plus
    24 bytes per row (12 read, 12 written) per partition pass that the path
    taken makes (0 for few groups, 1 for the key-range path); the bound with
    no partition pass is reported beside it."""
    import ctypes as C
    from pimdal_amd import datagen, _lib
    lo, hi = AGGR_SHAPES[shape]
    key, val = datagen.aggr_workload(ctx, n, lo, hi)
    cap = min(n, hi - lo + 1)
    ok, osum, ocnt = ctx.empty(cap, "uint32"), ctx.empty(cap, "int64"), ctx.empty(cap, "int64")
    g = C.c_uint64()
    fn = getattr(_lib.load(), f"pimdal_{which}_aggr_dev")

    def step():
        _lib.check(fn(ctx.handle, key.vp, val.vp, n, ok.vp, osum.vp, ocnt.vp, cap, C.byref(g)))

    tm = timed_steps(ctx, step, steps, warmup)
    # partition passes per step: the key-range scatter, or the hashed onesweep
    # passes, or the radix sort's passes of the sorted-runs path
    passes = sum(ctx.kernel_stat(k)[0] for k in ("ad_scatter", "part_onesweep")) / steps
    if ctx.kernel_stat("as_runs")[0]:
        passes += ctx.kernel_stat("radix_onesweep")[0] / steps
    p = round(passes)
    alg0 = 12 * n + 20 * g.value
    alg = alg0 + 24 * p * n
    res = {"rows": n, "groups": g.value, "key_range": [lo, hi], "s_per_step": tm, "rows_per_s": n / tm,
           "partition_passes": p, "alg_bytes_no_pass": alg0, "hbm_frac_no_pass": alg0 / tm / HBM_PEAK,
           "alg_bytes": alg, "alg_GBps": alg / tm / 1e9, "hbm_frac": alg / tm / HBM_PEAK,
           "kernels": kernel_breakdown(ctx, steps)}
    for a in (key, val, ok, osum, ocnt):
        a.free()
    return res


def bench_serve(ctx, nq=10_000, warm=300, n_fact=60_000_000, n_dim=15_000_000, rate_qps=1e6, slots=4,
                sweep=(2e3, 1e4, 1e5)):
    """Mixed select/join/group-by query stream over resident tables (6e7-row
    fact table, 1.5e7-row dimension table), batched vs one at a time, at the
    headline arrival rate and over a sweep of Poisson arrival rates (at most
    about 4 s of arrivals per rate)."""
    import numpy as np
    from pimdal_amd import datagen, serving
    budget_before = ctx.mem_stats()["budget"]
    ctx.set_budget(64 << 30)   # the serving stream runs under a fixed 64 GiB HBM budget
    fact = datagen.fact_table(ctx, n_fact, groups=n_dim)
    dim = datagen.dim_table(ctx, n_dim)
    srv = serving.Server(ctx, fact, dim, slots=slots, heap_bytes=4 << 30)
    wq = serving.gen_queries(warm, n_fact, rate_qps, seed=0x5EED0105)
    srv.run(wq, serving.BATCHED)
    srv.run(wq[:50], serving.SEQUENTIAL)

    def at(rate, n):
        q = serving.gen_queries(n, n_fact, rate, seed=0x5EED0005)
        r_seq, st_seq = srv.run(q, serving.SEQUENTIAL)
        r_bat, st_bat = srv.run(q, serving.BATCHED)
        same = bool(np.array_equal(r_seq["count"], r_bat["count"]) and np.array_equal(r_seq["checksum"], r_bat["checksum"]))
        return {"queries": n, "arrival_rate_qps": rate,
                "batched": {k: st_bat[k] for k in ("qps", "p50_us", "p99_us", "mean_us", "batches", "max_inflight",
                                                   "heap_peak_bytes")},
                "sequential": {k: st_seq[k] for k in ("qps", "p50_us", "p99_us", "mean_us")},
                "qps_speedup": st_bat["qps"] / st_seq["qps"] if st_seq["qps"] else None, "results_identical": same}

    res = at(rate_qps, nq)
    res["slots"] = slots
    res["rate_sweep"] = [at(r, int(min(nq, 4 * r))) for r in sweep]
    res["hbm_budget_bytes"] = 64 << 30
    res["hbm_live_bytes"] = ctx.mem_stats()["live"]   # resident tables, server indexes and the serving heap
    res["heap_bytes"] = 4 << 30
    srv.close()
    for d in (fact, dim):
        for a in d.values():
            a.free()
    ctx.set_budget(budget_before)
    return res


# ---------------- CPU baseline: the oracle's restatement on this host ----------------
#
# Every operator line above is timed again on the host CPU, on a bounded
# sample of the same seeded workload (host generators of the same recipes),
# at the two thread settings of the reference's CPU runs: all cores, and 16
# OpenMP threads on NUMA node 0 (reference/db_operators/profile.sh:6-13 runs
# each operator under `numactl --cpunodebind=0 --membind=0` with
# OMP_NUM_THREADS=16).  Each setting runs in a child process of this script
# (`--cpu-child`, which never touches the GPU), started before the GPU is
# initialised, so that the thread count and CPU binding hold for every
# thread.  Per operator: inputs generated (not timed), then the median of
# CPU_REPS timed runs; output allocation is inside the timing, as in
# reference/db_operators/select.cpp:115-135.

CPU_REPS = 5


def _cpu_cases():
    """(operator line, sample, rows counted, inputs factory, run) for every
    operator; rows are counted as the GPU line counts them."""
    import numpy as np
    from oracle import orc
    from pimdal_amd import datagen   # host-side generators only (Q3 tables, packed dates)

    def bloom_inputs(nb, npr):
        bk = orc.gen_u32(0xB10, 1, nb)
        pk, pp = orc.gen_u32(0xB10, 2, npr), orc.gen_u32(0xB10, 3, npr)
        for q in range(0, npr, 4 * nb):
            m = min(nb, npr - q)
            pk[q:q + m] = bk[:m]
        return bk, pk, pp

    def bloom_run(bits):
        def run(bk, pk, pp):
            f = orc.bloom_build(bk, bits, 32)
            return orc.bloom_probe(pk, pp, bits, 32, f)
        return run

    q3_date = int(datagen.pack_days(np.array([9204]))[0])
    q3_seg = datagen.Q3_SEGMENTS.index("BUILDING")
    q4_lo, q4_hi = (int(x) for x in datagen.pack_days(np.array([8582, 8674])))
    q5_lo, q5_hi = (int(x) for x in datagen.pack_days(np.array([8766, 9131])))
    cases = [
        ("sort_2^28", "2^24 uniform (u32 key, u32 row id) pairs, stable sort", 1 << 24,
         lambda: orc.sort_workload(1 << 24), orc.sort_pairs),
        ("sort_zipf_2^28", "2^24 pairs, keys fmix32(Zipf 1.0 over 2^24 values), stable sort", 1 << 24,
         lambda: orc.sort_workload(1 << 24, zipf=True), orc.sort_pairs),
        ("select_2^27", "2^27 (u32 key, u32 payload) rows (the whole input), key < 429496730 (10 %)", 1 << 27,
         lambda: orc.select_workload(1 << 27), lambda k, v: orc.select(k, v, 0, 429496730)),
    ]
    for which, fn in (("hash", orc.hash_join), ("smj", orc.sort_merge_join)):
        for z in (False, True):
            cases.append((f"{which}_join{'_zipf' if z else ''}",
                          f"build 2^21 unique x probe 2^24 {'Zipf 1.0' if z else 'uniform'} FK (one eighth of each side)",
                          (1 << 21) + (1 << 24), (lambda z=z: orc.join_workload(1 << 21, 1 << 24, zipf=z)), fn))
    for nb, bits in ((1 << 16, 5), (1 << 20, 9)):
        cases.append((f"bloom_semi_join_2^{nb.bit_length() - 1}x2^27",
                      f"2^{nb.bit_length() - 1} build keys into {1 << bits} filters, probe 2^24 rows",
                      nb + (1 << 24), (lambda nb=nb: bloom_inputs(nb, 1 << 24)), bloom_run(bits)))
    # aggregation: the whole 2^27 rows where the oracle is fast, else the
    # first rows of the same draw
    for which, shape, rows in (("hash", "k50", 1 << 27), ("hash", "k4000", 1 << 27), ("hash", "k2^24", 1 << 20),
                               ("sort", "k25", 1 << 24), ("sort", "k2^24", 1 << 24)):
        lo, hi = AGGR_SHAPES[shape]
        what = "the whole input" if rows == 1 << 27 else "the first rows of the 2^27-row draw"
        cases.append((f"{which}_aggr_2^27_{shape}", f"2^{rows.bit_length() - 1} rows ({what}), keys uniform on "
                      f"[{lo}, {hi}]", rows, (lambda lo=lo, hi=hi, rows=rows: orc.aggr_workload(rows, lo, hi)),
                      orc.hash_aggr if which == "hash" else orc.sort_aggr))
    fact = lambda: (orc.fact_table(60_000_000, groups=15_000_000),)
    cases += [
        ("q1_flags", "the whole 6e7-row fact table, date <= 2505, groups by (flag1, flag2)", 60_000_000, fact,
         lambda t: orc.q1_flags(t, 2505)),
        ("q1_orderkey", "2^20-row fact table with 2^18 orderkeys (4 rows per key as in 6e7 rows / 1.5e7 keys), "
         "date <= 2505", 1 << 20, lambda: (orc.fact_table(1 << 20, groups=1 << 18),), lambda t: orc.q1_key(t, 2505)),
        ("q6", "the whole 6e7-row fact table, TPC-H Q6 predicate", 60_000_000, fact, lambda t: orc.q6(t)),
        ("q3_sf10", "7.5e6 lineitems, 1.875e6 orders, 1.875e5 customers (SF 1.25), top 10", 7_500_000 + 1_875_000 + 187_500,
         lambda: (datagen.q3_tables(7_500_000),), lambda t: orc.q3(t, q3_seg, q3_date, cap=10)),
        ("q4_sf10", "1.5e7 lineitems, 3.75e6 orders (SF 2.5), orders of 1993-07-01 .. 1993-09-30",
         15_000_000 + 3_750_000, lambda: (datagen.q4_tables(15_000_000),), lambda t: orc.q4(t, q4_lo, q4_hi)),
        ("q5_sf10", "1.5e7 lineitems, 3.75e6 orders, 3.75e5 customers, 2.5e4 suppliers (SF 2.5), region ASIA, "
         "orders of 1994", 15_000_000 + 3_750_000 + 375_000 + 25_000, lambda: (datagen.q5_tables(15_000_000),),
         lambda t: orc.q5(t, datagen.Q5_REGIONS.index("ASIA"), q5_lo, q5_hi)),
        ("date32_pack_6e7", "2^20 date32 values over 1992-1998 (the C library's gmtime_r per value, which "
         "takes glibc's time-zone lock)", 1 << 20,
         lambda: (orc.gen_range(np.int32, 0xDA7E, 1, 1 << 20, 8035, 8035 + 2556),), orc.date32_pack),
    ]
    return cases


def cpu_child(threads):
    """Times every case at `threads` OpenMP threads; prints one JSON object."""
    from oracle import orc
    orc.set_threads(threads)
    out = {}
    for name, sample, rows, make, run in _cpu_cases():
        args = make()
        if not isinstance(args, tuple):
            args = (args,)
        times = []
        for _ in range(CPU_REPS):
            t0 = time.perf_counter()
            run(*args)
            times.append(time.perf_counter() - t0)
        del args
        t = statistics.median(times)
        out[name] = {"rows_per_s": rows / t, "median_s": t, "min_s": min(times), "max_s": max(times),
                     "rows": rows, "sample": sample}
        log(f"cpu baseline ({threads} threads): {name} {rows / t / 1e6:.1f} M rows/s")
    print(json.dumps({"threads": orc.num_threads(), "cases": out}), flush=True)


def _cpulist(text):
    cpus = set()
    for part in text.strip().split(","):
        if "-" in part:
            a, b = part.split("-")
            cpus.update(range(int(a), int(b) + 1))
        elif part:
            cpus.add(int(part))
    return cpus


def host_info():
    """CPU model, NUMA nodes with their CPUs, this process's CPU affinity and
    the cgroup CPU quota."""
    info = {"cpu_model": None, "numa_nodes": {}, "affinity_cpus": len(os.sched_getaffinity(0)),
            "online_cpus": os.cpu_count(), "cgroup_cpu_quota": None}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    for d in sorted(Path("/sys/devices/system/node").glob("node[0-9]*")):
        try:
            info["numa_nodes"][d.name] = (d / "cpulist").read_text().strip()
        except OSError:
            pass
    try:
        q, per = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if q != "max":
            info["cgroup_cpu_quota"] = int(q) / int(per)
    except (OSError, ValueError):
        pass
    return info


def cpu_baselines(timeout_s=300):
    """Runs the CPU child at both thread settings; returns the cpu_baseline
    object: the headline sort at all cores as value, every operator at both
    settings, and the host description."""
    import math
    import shutil
    import subprocess
    info = host_info()
    allowed = os.sched_getaffinity(0)
    n_all = len(allowed)
    if info["cgroup_cpu_quota"]:
        n_all = max(1, min(n_all, math.floor(info["cgroup_cpu_quota"])))
    node0 = _cpulist(info["numa_nodes"].get("node0", "")) & allowed or allowed
    numactl = shutil.which("numactl")
    settings = {"all_cores": (n_all, None), "node0_16": (16, node0)}
    res = {}
    for name, (threads, cpus) in settings.items():
        cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-child", str(threads)]
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        pre = None
        if cpus is not None:
            if numactl:
                cmd = [numactl, "--cpunodebind=0", "--membind=0"] + cmd
                binding = "numactl --cpunodebind=0 --membind=0"
            else:
                pre = lambda cpus=cpus: os.sched_setaffinity(0, cpus)
                binding = (f"affinity to node0's {len(cpus)} CPUs (no numactl on this host: memory on node 0 "
                           f"by first touch from threads bound there)")
        else:
            binding = f"none ({threads} threads over the {len(allowed)} CPUs this process may use)"
        log(f"bench: cpu baseline, {name}")
        p = subprocess.run(cmd, env=env, preexec_fn=pre, stdout=subprocess.PIPE, timeout=timeout_s, check=True)
        r = json.loads(p.stdout.decode().strip().splitlines()[-1])
        res[name] = {"threads": r["threads"], "binding": binding, "cases": r["cases"]}
    head = res["all_cores"]["cases"]["sort_2^28"]
    return {"value": head["rows_per_s"], "unit": "rows/s", "cores": res["all_cores"]["threads"], "kind": "port",
            "sample": head["sample"] + f", median of {CPU_REPS}", "reps": CPU_REPS, "host": info,
            "settings": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--only", default="", help="comma-separated operator groups (zipf, sortselect, join, bloom, "
                    "groupby, aggr, q3, q4, ingest, serve); the 2^28 sort always runs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="append", default=[], help="tuning knob key=value (repeatable)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one 1 GiB device-to-device copy (the profiler's byte-counter reference)")
    ap.add_argument("--cpu-child", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args.cpu_child)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; "
              "run with --gpus 1 (how the operators would shard is documented with the design)", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline runs first, in child processes, before this process
    # initialises the GPU
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines()

    only = {o for o in args.only.split(",") if o}
    sel = lambda *names: not only or "all" in only or any(n in only for n in names)
    import pimdal_amd
    ctx = pimdal_amd.Context(0)
    for kv in args.tune:
        k, v = kv.split("=")
        ctx.set_tune(k, int(v))
    ops_out = {}
    if args.calibrate:
        import ctypes as C
        from pimdal_amd._lib import call
        a, b = ctx.empty(1 << 28, "uint32"), ctx.empty(1 << 28, "uint32")
        call("pimdal_dev_copy", ctx.handle, b.vp, a.vp, 1 << 30)
        ctx.sync()
        a.free(); b.free()
    log("bench: sort 2^28 uniform")
    srt = bench_sort(ctx, args.steps, args.warmup)
    if sel("zipf"):
        log("bench: sort 2^28 zipf")
        ops_out["sort_zipf_2^28"] = bench_sort(ctx, max(3, args.steps // 2), 1, zipf=True)
    if sel("sortselect"):
        log("bench: select 2^27")
        ops_out["select_2^27"] = bench_select(ctx, args.steps, args.warmup)
    if sel("join"):
        for which in ("hash", "smj"):
            for z in (False, True):
                log(f"bench: {which} join 2^24 x 2^27 zipf={z}")
                ops_out[f"{which}_join{'_zipf' if z else ''}"] = bench_join(ctx, max(3, args.steps // 2), 1, which, z)
        # small build side: the global-table mode against the partitioned one
        for mode in (1, 0):
            log(f"bench: hash join 2^20 x 2^27 hj_mode={mode}")
            ctx.set_tune("hj_mode", mode)
            ops_out[f"hash_join_build2^20_mode{mode}"] = bench_join(ctx, max(3, args.steps // 2), 1, "hash", False,
                                                                    nr=1 << 20)
        ctx.set_tune("hj_mode", -1)
    if sel("bloom"):
        for nb, bits in ((1 << 16, 5), (1 << 20, 9)):
            log(f"bench: bloom semi-join 2^{nb.bit_length() - 1} x 2^27, {1 << bits} filters")
            ops_out[f"bloom_semi_join_2^{nb.bit_length() - 1}x2^27"] = bench_bloom(ctx, args.steps, args.warmup, nb, bits)
    if sel("groupby"):
        for by_key in (False, True):
            log(f"bench: filter+group-by 6e7 by_key={by_key}")
            ops_out["q1_orderkey" if by_key else "q1_flags"] = bench_q1(ctx, max(3, args.steps // 2), 1, by_key)
            log(f"bench: filter+group-by 6e7 by_key={by_key}, one captured plan")
            ops_out["q1_orderkey_plan" if by_key else "q1_flags_plan"] = bench_q1_plan(ctx, max(3, args.steps // 2), 1,
                                                                                         by_key)
        log("bench: TPC-H Q6 shape 6e7")
        ops_out["q6"] = bench_q6(ctx, max(3, args.steps // 2), 1)
    if sel("aggr"):
        for which, shape in (("hash", "k50"), ("hash", "k4000"), ("hash", "k2^24"), ("sort", "k25"),
                             ("sort", "k2^24")):
            log(f"bench: {which} aggregation 2^27 keys {AGGR_SHAPES[shape]}")
            ops_out[f"{which}_aggr_2^27_{shape}"] = bench_aggr(ctx, max(3, args.steps // 2), 1, which, shape)
    if sel("q3"):
        log("bench: TPC-H Q3 shape, 6e7 lineitems")
        ops_out["q3_sf10"] = bench_q3(ctx, max(3, args.steps // 2), 1)
    if sel("q3", "q4", "q5"):
        log("bench: TPC-H Q4 shape, 6e7 lineitems")
        ops_out["q4_sf10"] = bench_q4(ctx, max(3, args.steps // 2), 1)
        log("bench: TPC-H Q5 shape, 6e7 lineitems")
        ops_out["q5_sf10"] = bench_q5(ctx, max(3, args.steps // 2), 1)
    if sel("ingest"):
        log("bench: date32 to packed dates 6e7")
        ops_out["date32_pack_6e7"] = bench_date32(ctx, args.steps, args.warmup)
    if sel("serve"):
        log("bench: serving mixed stream")
        ops_out["serve_mixed_stream"] = bench_serve(ctx)
        # the same stream with batched joins probing the key index per row
        # (no resident join index), to separate batching from the index
        log("bench: serving mixed stream, probe index")
        ctx.set_tune("serve_join_index", 0)
        ops_out["serve_mixed_stream_probe_index"] = bench_serve(ctx, sweep=())
        ctx.set_tune("serve_join_index", 1)
    if cpu:
        for name, st in cpu["settings"].items():
            for op, r in st["cases"].items():
                tgt = ops_out.get(op) if op != "sort_2^28" else srt
                if tgt is not None and tgt.get("rows_per_s"):
                    tgt.setdefault("cpu_rows_per_s", {})[name] = r["rows_per_s"]
                    tgt.setdefault("gpu_over_cpu", {})[name] = tgt["rows_per_s"] / r["rows_per_s"]
    pass_s = srt["pass_ms"] / 1e3
    achieved = srt["pass_alg_bytes"] / pass_s if pass_s else None
    line = {
        "metric": METRIC,
        "value": srt["rows_per_s"],
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": srt["s_per_step"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "stable radix sort of 2^28 (u32 key, u32 payload) pairs, keys uniform on [0,2^32)",
                   "rows": srt["n"], "hbm_trips": srt["hbm_trips"], "digit_bits": 8,
                   "scheme": ("onesweep passes over the two high digits, then one pass sorting each bucket of "
                              "equal high bits by the two low digits in LDS") if srt["bucket_launches"]
                   else "one onesweep pass per digit",
                   "alg_bytes_per_row": 4 + 16 * srt["hbm_trips"], "sort_alg_GBps": srt["alg_GBps"],
                   "sort_hbm_frac": srt["hbm_frac"], "hist_ms": srt["hist_ms"], "bucket_ms": srt["bucket_ms"],
                   "count_ms": srt["count_ms"], "tscan_ms": srt["tscan_ms"],
                   "keys_payload_GBps": srt["keys_payload_GBps"], "gpu_over_cpu": srt.get("gpu_over_cpu")},
        "roofline": {"kernel": "radix_onesweep", "bound": "hbm",
                     "achieved": achieved / 1e9 if achieved else None, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK if achieved else None,
                     "traffic": load_traffic("radix_onesweep_2p28"),
                     "alg_bytes_per_launch": srt["pass_alg_bytes"], "avg_launch_ms": srt["pass_ms"]},
        "cpu_baseline": cpu,
        "operators": ops_out,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
