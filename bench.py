#!/usr/bin/env python3
"""Benchmark of the DB-operator hot path on one MI355X.

Headline workload: stable LSD radix sort of 2^28 (uint32 key, uint32 payload)
pairs with keys uniform on [0, 2^32).  One step = one sort of the whole
batch, inputs already resident in HBM.  The JSON line also carries the other
operators' rows/s at their configured sizes, the dominant kernel's roofline
(hipEvent time on the stream it runs on), and the CPU restatement timed on
a bounded sample on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--only sort|zipf|sortselect|join|bloom|groupby|q3|ingest|serve]
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
    for a in (keys, vals, ok, ov):
        a.free()
    return {
        "n": n, "s_per_step": t, "rows_per_s": n / t,
        "alg_bytes": 68 * n, "alg_GBps": 68 * n / t / 1e9, "hbm_frac": 68 * n / t / HBM_PEAK,
        "keys_payload_GBps": 8 * n / t / 1e9,
        "pass_ms": pass_ms, "pass_launches": launches, "hist_ms": hist_ms,
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


KERNELS = ("bloom_build", "bloom_flags", "date32_pack", "part_hist", "part_count", "part_tscan", "radix_count", "radix_tscan", "part_onesweep", "part_bounds", "scan", "hj_chunks", "hj_join", "hj_join_count",
           "hj_join_write", "hj_compact", "hg_build", "hg_probe", "hg_probe_count", "hg_probe_write", "radix_hist",
           "radix_scan", "radix_onesweep", "smj_dupcheck", "smj_ranges", "smj_merge", "smj_expand", "q1_domain",
           "q1_lut", "q1_dense", "q1_reduce", "q1_final", "gb_direct", "gb_direct_emit", "gb_chunks", "gb_hist",
           "gb_scatter", "gb_lds", "gb_global", "gb_emit_global", "q1_unpack", "agg_heads", "agg_run_reduce",
           "select", "select_count", "select_write", "q1_pack", "q6", "q6_final", "smj_pack", "fill_u32", "fill_u64",
           "q3_cust", "q3_osel", "q3_orders", "q3_lineitem", "q3_emit", "q3_order")


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


def cpu_baseline_sort(sample_n=1 << 24, reps=3):
    """The CPU restatement (oracle/) of the stable pair sort, all host threads."""
    import numpy as np
    from oracle import orc
    k, v = orc.sort_workload(sample_n)
    orc.sort_pairs(k[:1024], v[:1024])
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        orc.sort_pairs(k, v)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    return {"value": sample_n / t, "unit": "rows/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"2^{int(np.log2(sample_n))} uniform (u32,u32) pairs, stable sort, median of {reps}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--only", default="")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="append", default=[], help="tuning knob key=value (repeatable)")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one 1 GiB device-to-device copy (the profiler's byte-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; "
              "run with --gpus 1 (how the operators would shard is documented with the design)", file=sys.stderr)
        sys.exit(2)

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
    if args.only in ("", "all", "zipf"):
        log("bench: sort 2^28 zipf")
        ops_out["sort_zipf_2^28"] = bench_sort(ctx, max(3, args.steps // 2), 1, zipf=True)
    if args.only in ("", "all", "sortselect"):
        log("bench: select 2^27")
        ops_out["select_2^27"] = bench_select(ctx, args.steps, args.warmup)
    if args.only in ("", "all", "join"):
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
    if args.only in ("", "all", "bloom"):
        for nb, bits in ((1 << 16, 5), (1 << 20, 9)):
            log(f"bench: bloom semi-join 2^{nb.bit_length() - 1} x 2^27, {1 << bits} filters")
            ops_out[f"bloom_semi_join_2^{nb.bit_length() - 1}x2^27"] = bench_bloom(ctx, args.steps, args.warmup, nb, bits)
    if args.only in ("", "all", "groupby"):
        for by_key in (False, True):
            log(f"bench: filter+group-by 6e7 by_key={by_key}")
            ops_out["q1_orderkey" if by_key else "q1_flags"] = bench_q1(ctx, max(3, args.steps // 2), 1, by_key)
            log(f"bench: filter+group-by 6e7 by_key={by_key}, one captured plan")
            ops_out["q1_orderkey_plan" if by_key else "q1_flags_plan"] = bench_q1_plan(ctx, max(3, args.steps // 2), 1,
                                                                                         by_key)
        log("bench: TPC-H Q6 shape 6e7")
        ops_out["q6"] = bench_q6(ctx, max(3, args.steps // 2), 1)
    if args.only in ("", "all", "q3"):
        log("bench: TPC-H Q3 shape, 6e7 lineitems")
        ops_out["q3_sf10"] = bench_q3(ctx, max(3, args.steps // 2), 1)
    if args.only in ("", "all", "ingest"):
        log("bench: date32 to packed dates 6e7")
        ops_out["date32_pack_6e7"] = bench_date32(ctx, args.steps, args.warmup)
    if args.only in ("", "all", "serve"):
        log("bench: serving mixed stream")
        ops_out["serve_mixed_stream"] = bench_serve(ctx)
        # the same stream with batched joins probing the key index per row
        # (no resident join index), to separate batching from the index
        log("bench: serving mixed stream, probe index")
        ctx.set_tune("serve_join_index", 0)
        ops_out["serve_mixed_stream_probe_index"] = bench_serve(ctx, sweep=())
        ctx.set_tune("serve_join_index", 1)
    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline")
        cpu = cpu_baseline_sort()
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
        "config": {"workload": "stable LSD radix sort of 2^28 (u32 key, u32 payload) pairs, keys uniform on [0,2^32)",
                   "rows": srt["n"], "passes": 4, "digit_bits": 8,
                   "alg_bytes_per_row": 68, "sort_alg_GBps": srt["alg_GBps"], "sort_hbm_frac": srt["hbm_frac"],
                   "hist_ms": srt["hist_ms"], "count_ms": srt["count_ms"], "tscan_ms": srt["tscan_ms"],
                   "keys_payload_GBps": srt["keys_payload_GBps"]},
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
