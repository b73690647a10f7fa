"""Radix sort on the GPU vs the CPU restatement (stable pair sort)."""
import numpy as np
import pytest

from pimdal_amd import operators as ops, datagen

pytestmark = pytest.mark.gpu


def gpu_sort(ctx, k, v, begin=-1, end=-1):
    dk = ctx.to_device(k)
    dv = ctx.to_device(v) if v is not None else None
    ok, ov = ops.sort_pairs(ctx, dk, dv, begin, end)
    return ok.numpy(), (ov.numpy() if ov is not None else None)


@pytest.mark.parametrize("n", [1, 2, 100, 8191, 8192, 8193, 1 << 20])
def test_sort_uniform_pairs(ctx, orc, n):
    k, v = orc.sort_workload(n, seed=900 + n)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_sort_zipf_pairs_stable(ctx, orc):
    k, v = orc.sort_workload(1 << 21, zipf=True)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_sort_keys_only(ctx, orc):
    k, _ = orc.sort_workload(300_000, seed=4)
    gk, _ = gpu_sort(ctx, k, None)
    assert np.array_equal(gk, orc.sort_keys(k))


def test_sort_u64_payload(ctx, orc):
    k = orc.gen_range(np.uint32, 6, 1, 200_000, 0, 1000)
    v = orc.gen_range(np.int64, 6, 2, 200_000, -(1 << 40), 1 << 40)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v.view(np.uint64))
    assert np.array_equal(gk, ek) and np.array_equal(gv.view(np.uint64), ev)


def test_sort_constant_and_narrow_keys(ctx, orc):
    n = 100_000
    v = np.arange(n, dtype=np.uint32)
    k = np.full(n, 7, np.uint32)                          # no varying bit: pure copy
    gk, gv = gpu_sort(ctx, k, v)
    assert np.array_equal(gk, k) and np.array_equal(gv, v)
    k = orc.gen_range(np.uint32, 2, 1, n, 0, (1 << 24) - 1)   # 24-bit keys: 3 passes
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


@pytest.mark.parametrize("begin,end", [(0, 32), (0, 24), (8, 16), (4, 13)])
def test_sort_explicit_bit_range(ctx, orc, begin, end):
    k, v = orc.sort_workload(70_000, seed=11)
    gk, gv = gpu_sort(ctx, k, v, begin, end)
    mk = (k >> np.uint32(begin)) & np.uint32((1 << (end - begin)) - 1)
    order = np.argsort(mk, kind="stable")
    assert np.array_equal(gv, v[order]) and np.array_equal(gk, k[order])


@pytest.mark.parametrize("zipf", [False, True])
def test_sort_2_28_properties(ctx, zipf):
    """Full size: sorted, stable, a permutation, and out_key[i] == in_key[out_row[i]]."""
    n = 1 << 28
    dk, dv = datagen.sort_workload(ctx, n, zipf=zipf)
    ctx.set_profiling(True)
    ctx.reset_stats()
    ok, ov = ops.sort_pairs(ctx, dk, dv)
    # uniform keys take the bucket scheme by default, Zipf keys plain passes
    assert ctx.kernel_stat("radix_bucket")[0] == (0 if zipf else 1)
    ctx.set_profiling(False)
    if not zipf:
        # the same sort by one global pass per digit gives the same arrays
        ctx.set_tune("sort_bucket", 0)
        try:
            pk, pv = ops.sort_pairs(ctx, dk, dv)
        finally:
            ctx.clear_tune("sort_bucket")
        assert np.array_equal(pk.numpy(), ok.numpy()) and np.array_equal(pv.numpy(), ov.numpy())
        pk.free()
        pv.free()
    k = dk.numpy()
    dk.free()
    dv.free()
    gk = ok.numpy()
    ok.free()
    gv = ov.numpy()
    ov.free()
    assert np.all(gk[:-1] <= gk[1:])
    seen = np.zeros(n, np.bool_)
    seen[gv] = True
    assert seen.all()
    del seen
    assert np.array_equal(k[gv], gk)
    ties = gk[:-1] == gk[1:]
    assert np.all(gv[:-1][ties] < gv[1:][ties])


@pytest.mark.parametrize("bits", [(-1, -1), (0, 32)])
def test_sort_2_28_oversized_buckets_count_low_digits_late(ctx, orc, bits):
    """At 2^28 the first read counts only the two high digits.  Keys whose
    high byte equals the second fill 256 buckets of about 2^20 rows: the
    bucket pass flags them, the low digits are counted then, and every digit
    is sorted by ordinary passes.  Row ids as payloads; sorted, stable, a
    permutation, out_key[i] == in_key[out_row[i]]."""
    n = 1 << 28
    b = orc.gen_range(np.uint32, 12, 1, n, 0, 255)
    k = (b << np.uint32(24)) | (b << np.uint32(16)) | orc.gen_range(np.uint32, 12, 2, n, 0, (1 << 16) - 1)
    del b
    ctx.set_profiling(True)
    ctx.reset_stats()
    try:
        dk = ctx.to_device(k)
        dv = ctx.to_device(np.arange(n, dtype=np.uint32))
        ok, ov = ops.sort_pairs(ctx, dk, dv, *bits)
        assert ctx.kernel_stat("radix_bucket")[0] == 1
        assert ctx.kernel_stat("radix_hist")[0] == 2 and ctx.kernel_stat("radix_onesweep")[0] == 2 + 4
    finally:
        ctx.set_profiling(False)
    dk.free()
    dv.free()
    gk = ok.numpy()
    ok.free()
    gv = ov.numpy()
    ov.free()
    assert np.all(gk[:-1] <= gk[1:])
    seen = np.zeros(n, np.bool_)
    seen[gv] = True
    assert seen.all()
    del seen
    assert np.array_equal(k[gv], gk)
    ties = gk[:-1] == gk[1:]
    assert np.all(gv[:-1][ties] < gv[1:][ties])


def test_autotune_sets_knobs_and_results_stay_exact(ctx, orc, tmp_path):
    cache = str(tmp_path / "tune.txt")
    ctx.autotune("sort", 1 << 22, cache)
    assert ctx.get_tune("sort_block") in (256, 512)
    ctx.load_tuning(cache)
    k, v = orc.sort_workload(1 << 20, seed=5)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)
    assert open(cache).read().split()[0] == "sort"


@pytest.fixture
def ballot_rank(ctx):
    """Onesweep ranking by per-bit ballots (sort_rank 1) instead of LDS
    match masks."""
    ctx.set_tune("sort_rank", 1)
    yield
    ctx.set_tune("sort_rank", 0)


@pytest.mark.parametrize("n", [1, 8193, 1 << 20])
def test_sort_ballot_rank_uniform(ctx, orc, ballot_rank, n):
    k, v = orc.sort_workload(n, seed=70 + n)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_sort_ballot_rank_zipf_and_partial_bits(ctx, orc, ballot_rank):
    k, v = orc.sort_workload(1 << 21, zipf=True)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)
    k, v = orc.sort_workload(70_000, seed=12)
    gk, gv = gpu_sort(ctx, k, v, 4, 13)
    mk = (k >> np.uint32(4)) & np.uint32((1 << 9) - 1)
    order = np.argsort(mk, kind="stable")
    assert np.array_equal(gv, v[order]) and np.array_equal(gk, k[order])


def test_sort_ballot_rank_u64_payload_and_hash_join(ctx, orc, ballot_rank):
    k = orc.gen_range(np.uint32, 6, 1, 200_000, 0, 1000)
    v = orc.gen_range(np.int64, 6, 2, 200_000, -(1 << 40), 1 << 40)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v.view(np.uint64))
    assert np.array_equal(gk, ek) and np.array_equal(gv.view(np.uint64), ev)
    # the hashed partition passes of the hash join run the same kernel
    rk, rp, sk, sp = orc.join_workload(1 << 17, 1 << 19)   # above the global-table size: partitioned
    op, ob, cnt = ops.hash_join(ctx, *[ctx.to_device(a) for a in (rk, rp, sk, sp)])
    ep, eb = orc.hash_join(rk, rp, sk, sp)
    got = np.sort((op.numpy(cnt).astype(np.uint64) << np.uint64(32)) | ob.numpy(cnt))
    exp = np.sort((ep.astype(np.uint64) << np.uint64(32)) | eb)
    assert np.array_equal(got, exp)


PASS_VARIANTS = {
    "window1": {"sort_lbw": 1},
    "window4": {"sort_lbw": 4},
    "window8": {"sort_lbw": 8},
    "window32": {"sort_lbw": 32},
    "split_lookback": {"sort_lb_split": 1},
    "split_lookback_window8": {"sort_lb_split": 1, "sort_lbw": 8},
    "nontemporal_loads": {"sort_nt": 1},
    "nontemporal_loads_stores": {"sort_nt": 3},
    "launch_order_tiles": {"sort_xcd_tiles": 0},
    "persistent_launch_refused_at_pass0": {"sort_xcd_refuse_pass": 0},
    "persistent_launch_refused_at_pass1": {"sort_xcd_refuse_pass": 1},
    "xcd_persistent_runs1": {"sort_xcd_tiles": 1},
    "xcd_persistent_runs5": {"sort_xcd_tiles": 5},
    "xcd_persistent_runs16_all_digit_widths": {"sort_xcd_tiles": 16},
    "block1024_items8": {"sort_block": 1024},
    "block512_items8": {"sort_items": 8},
    "block256_early": {"sort_block": 256, "sort_early_lookback": 1},
    "xcd_persistent_block256": {"sort_block": 256},
    "block256_refused_at_pass1": {"sort_block": 256, "sort_xcd_refuse_pass": 1},
    "ballot_rank": {"sort_rank": 1},
    "counted_xcd_order": {"sort_scheme": 1},
    "counted_ticket_order": {"sort_scheme": 1, "sort_rts_order": 1},
    "counted_block1024": {"sort_scheme": 1, "sort_block": 1024},
    "plain_passes_no_bucket": {"sort_bucket": 0},
    "bucket_after_launch_order_tiles": {"sort_xcd_tiles": 0, "sort_bucket": 2},
    "bucket_scheme": {"sort_bucket": 2},
}


@pytest.fixture(params=sorted(PASS_VARIANTS))
def pass_variant(ctx, request):
    """Pass variants: the chained pass (decoupled look-back) with other
    look-back windows, block shapes and rankings; the counted
    (reduce-then-scan) pass in its tile orders and shapes."""
    knobs = PASS_VARIANTS[request.param]
    for k, v in knobs.items():
        ctx.set_tune(k, v)
    yield request.param
    for k in knobs:
        ctx.clear_tune(k)


@pytest.mark.parametrize("n", [1, 8191, 8193, 1 << 20])
def test_sort_pass_variant_uniform(ctx, orc, pass_variant, n):
    k, v = orc.sort_workload(n, seed=170 + n)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_sort_pass_variant_zipf_bits_payloads_and_hash_join(ctx, orc, pass_variant):
    k, v = orc.sort_workload(1 << 21, zipf=True)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)
    k, v = orc.sort_workload(70_000, seed=13)
    gk, gv = gpu_sort(ctx, k, v, 4, 13)
    mk = (k >> np.uint32(4)) & np.uint32((1 << 9) - 1)
    order = np.argsort(mk, kind="stable")
    assert np.array_equal(gv, v[order]) and np.array_equal(gk, k[order])
    # 24-bit keys (top digit constant), keys only, 8-byte payloads
    k = orc.gen_range(np.uint32, 2, 1, 100_000, 0, (1 << 24) - 1)
    gk, _ = gpu_sort(ctx, k, None)
    assert np.array_equal(gk, orc.sort_keys(k))
    k = (orc.gen_range(np.uint32, 3, 1, 100_000, 0, 255) << np.uint32(8)).astype(np.uint32)  # lowest digit constant
    v = orc.gen_range(np.int64, 6, 2, 100_000, -(1 << 40), 1 << 40)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v.view(np.uint64))
    assert np.array_equal(gk, ek) and np.array_equal(gv.view(np.uint64), ev)
    rk, rp, sk, sp = orc.join_workload(1 << 17, 1 << 19)   # partitioned: the hashed passes
    op, ob, cnt = ops.hash_join(ctx, *[ctx.to_device(a) for a in (rk, rp, sk, sp)])
    ep, eb = orc.hash_join(rk, rp, sk, sp)
    got = np.sort((op.numpy(cnt).astype(np.uint64) << np.uint64(32)) | ob.numpy(cnt))
    exp = np.sort((ep.astype(np.uint64) << np.uint64(32)) | eb)
    assert np.array_equal(got, exp)


def test_sort_in_a_child_process_after_this_one_ran_persistent_passes(ctx, orc):
    """This process has run co-resident (cooperative) passes; a second process
    on the same GPU sorts too.  Where the runtime refuses that process's
    cooperative grid, its passes take one block per tile and stay exact.
    The child reports which path its passes took (the fallback launches the
    digit-offset scan that the co-resident passes do themselves); the
    exactness check holds for either."""
    import subprocess
    import sys
    from pathlib import Path
    k, v = orc.sort_workload(1 << 20, seed=91)
    gk, gv = gpu_sort(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)
    child = (
        "import numpy as np, pimdal_amd\n"
        "from pimdal_amd import operators as ops\n"
        "from oracle import orc\n"
        "k, v = orc.sort_workload(1 << 20, seed=92)\n"
        "with pimdal_amd.Context(0) as c:\n"
        "    c.set_profiling(True)\n"
        "    ok, ov = ops.sort_pairs(c, c.to_device(k), c.to_device(v))\n"
        "    ek, ev = orc.sort_pairs(k, v)\n"
        "    assert np.array_equal(ok.numpy(), ek) and np.array_equal(ov.numpy(), ev)\n"
        "    n_scan = c.kernel_stat('radix_scan')[0]\n"
        "    n_pass = c.kernel_stat('radix_onesweep')[0]\n"
        "    assert n_pass >= 1\n"
        "    print('child path:', 'one block per tile' if n_scan else 'co-resident', flush=True)\n"
        "print('child sort ok', flush=True)\n")
    root = Path(__file__).resolve().parents[1]
    res = subprocess.run([sys.executable, "-c", child], cwd=root, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "child sort ok" in res.stdout, res.stderr[-2000:]
    assert "child path:" in res.stdout
    print(res.stdout.strip().splitlines()[0])


# ---- the bucket scheme: global passes over the two high digits, then each
# bucket of equal high bits sorted by the low digits in LDS ----

def _sort_stats(ctx, k, v, begin=-1, end=-1, bucket=2):
    """Sort with sort_bucket = `bucket` (2: the bucket scheme wherever its
    estimate fits, at sizes below the ones it is chosen for by default);
    returns the bucket-pass and onesweep launch counts beside the result."""
    ctx.set_profiling(True)
    ctx.reset_stats()
    ctx.set_tune("sort_bucket", bucket)
    try:
        gk, gv = gpu_sort(ctx, k, v, begin, end)
        return gk, gv, ctx.kernel_stat("radix_bucket")[0], ctx.kernel_stat("radix_onesweep")[0]
    finally:
        ctx.clear_tune("sort_bucket")
        ctx.set_profiling(False)


@pytest.mark.parametrize("n", [1 << 20, (1 << 21) + 77, 1 << 22])
def test_sort_bucket_scheme_uniform(ctx, orc, n):
    k, v = orc.sort_workload(n, seed=300 + n)
    gk, gv, nb, npass = _sort_stats(ctx, k, v)
    assert nb == 1 and npass == 2, (nb, npass)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_sort_bucket_scheme_keys_only_and_u64_payload(ctx, orc):
    k, _ = orc.sort_workload(1 << 21, seed=31)
    gk, _, nb, _ = _sort_stats(ctx, k, None)
    assert nb == 1 and np.array_equal(gk, orc.sort_keys(k))
    v = orc.gen_range(np.int64, 6, 3, 1 << 21, -(1 << 40), 1 << 40)
    gk, gv, nb, _ = _sort_stats(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v.view(np.uint64))
    assert nb == 1 and np.array_equal(gk, ek) and np.array_equal(gv.view(np.uint64), ev)


def test_sort_bucket_scheme_24_bit_keys_and_bit_ranges(ctx, orc):
    n = (1 << 21) + 5
    v = np.arange(n, dtype=np.uint32)
    k = orc.gen_range(np.uint32, 2, 7, n, 0, (1 << 24) - 1)   # three digits: one low digit in LDS
    gk, gv, nb, npass = _sort_stats(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert nb == 1 and npass == 2 and np.array_equal(gk, ek) and np.array_equal(gv, ev)
    k, _ = orc.sort_workload(n, seed=33)
    for begin, end in ((0, 28), (3, 29), (0, 32)):
        gk, gv, nb, _ = _sort_stats(ctx, k, v, begin, end)
        mk = (k >> np.uint32(begin)) & np.uint32((1 << (end - begin)) - 1)
        order = np.argsort(mk, kind="stable")
        assert nb == 1 and np.array_equal(gv, v[order]) and np.array_equal(gk, k[order]), (begin, end)


def test_sort_bucket_scheme_duplicate_keys_stay_stable(ctx, orc):
    """2^16 distinct keys spread over 32 bits: every bucket holds runs of
    equal keys, whose payloads (row ids) must keep input order."""
    n = 1 << 21
    r = orc.gen_range(np.uint32, 4, 1, n, 0, (1 << 16) - 1).astype(np.uint64)
    k = ((r * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    v = np.arange(n, dtype=np.uint32)
    gk, gv, nb, _ = _sort_stats(ctx, k, v)
    ek, ev = orc.sort_pairs(k, v)
    assert nb == 1 and np.array_equal(gk, ek) and np.array_equal(gv, ev)


@pytest.mark.parametrize("counting", [0, 1])
def test_sort_bucket_counting_form_matches_ranking(ctx, orc, counting):
    """Four varying digits take the counting form of the bucket pass (one LDS
    atomic per row into 2048 coarse bins, each bin ordered by (low bits,
    row)); knob sort_bucket_count=0 takes the two ranking rounds.  Both
    against the oracle, with 2^16 repeated keys so equal keys share bins."""
    n = (1 << 22) + 333
    k, _ = orc.sort_workload(n, seed=41)
    k[::7] = k[: (n + 6) // 7][::-1]        # repeats, met out of order
    v = np.arange(n, dtype=np.uint32)
    ctx.set_tune("sort_bucket_count", counting)
    try:
        gk, gv, nb, npass = _sort_stats(ctx, k, v)
    finally:
        ctx.clear_tune("sort_bucket_count")
    ek, ev = orc.sort_pairs(k, v)
    assert nb == 1 and npass == 2 and np.array_equal(gk, ek) and np.array_equal(gv, ev)


@pytest.mark.parametrize("payload", ["u32", "none", "u64"])
def test_sort_bucket_counting_large_bins_go_to_ranking(ctx, orc, payload):
    """Buckets whose rows share their low 16 bits put more than 32 rows in
    one coarse bin: those chunks are handed, untouched, to the ranking
    kernel (radix_bucket_retry); the others stay in the counting form."""
    n = 1 << 22
    k, _ = orc.sort_workload(n, seed=43)
    hot = orc.gen_range(np.uint32, 9, 1, 300, 0, 0xFFFFFFFF)
    k[1::5] = hot[np.arange(k[1::5].size) % hot.size]      # 300 keys about 2800 times each
    lo = np.uint32(0xBEEF)
    k[2::11] = (k[2::11] & np.uint32(0xFFFF0000)) | lo    # equal low bits across buckets
    v = {"u32": np.arange(n, dtype=np.uint32), "none": None,
         "u64": orc.gen_range(np.int64, 9, 2, n, -(1 << 50), 1 << 50).view(np.uint64)}[payload]
    ctx.set_profiling(True)
    ctx.reset_stats()
    ctx.set_tune("sort_bucket", 2)
    try:
        gk, gv = gpu_sort(ctx, k, v)
        assert ctx.kernel_stat("radix_bucket")[0] == 1 and ctx.kernel_stat("radix_bucket_retry")[0] == 1
    finally:
        ctx.clear_tune("sort_bucket")
        ctx.set_profiling(False)
    if v is None:
        assert np.array_equal(gk, orc.sort_keys(k))
    else:
        ek, ev = orc.sort_pairs(k, v)
        assert np.array_equal(gk, ek) and np.array_equal(gv.view(ev.dtype), ev)


def test_sort_bucket_scheme_oversized_bucket_falls_back(ctx, orc):
    """High byte == second byte: the histograms look uniform but only 256
    buckets hold rows (16384 each, beyond a block's 8192), so the bucket pass
    flags them and every digit is sorted again by ordinary passes."""
    n = 1 << 22
    b = orc.gen_range(np.uint32, 8, 1, n, 0, 255)
    lo = orc.gen_range(np.uint32, 8, 2, n, 0, (1 << 16) - 1)
    k = ((b << np.uint32(24)) | (b << np.uint32(16)) | lo).astype(np.uint32)
    v = np.arange(n, dtype=np.uint32)
    gk, gv, nb, npass = _sort_stats(ctx, k, v)
    assert nb == 1 and npass == 2 + 4, (nb, npass)
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


def test_sort_bucket_scheme_off_small_and_zipf_take_plain_passes(ctx, orc):
    k, v = orc.sort_workload(1 << 21, zipf=True)
    gk, gv, nb, npass = _sort_stats(ctx, k, v)   # the estimate does not fit
    assert nb == 0 and npass == 4
    ek, ev = orc.sort_pairs(k, v)
    assert np.array_equal(gk, ek) and np.array_equal(gv, ev)
    k, v = orc.sort_workload(1 << 21, seed=35)
    ek, ev = orc.sort_pairs(k, v)
    for bucket in (0, 1):   # off; by default buckets of 32 rows are too small
        gk, gv, nb, npass = _sort_stats(ctx, k, v, bucket=bucket)
        assert nb == 0 and npass == 4 and np.array_equal(gk, ek) and np.array_equal(gv, ev)
