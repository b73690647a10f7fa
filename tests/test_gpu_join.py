"""Hash join and sort-merge join on the GPU vs the CPU restatement."""
import numpy as np
import pytest

from pimdal_amd import operators as ops, datagen
from tests.util import canon, same_multiset

pytestmark = pytest.mark.gpu


def run(ctx, fn, rk, rp, sk, sp):
    d = [ctx.to_device(a) for a in (rk, rp, sk, sp)]
    op, ob, k = fn(ctx, *d)
    return op.numpy(k), ob.numpy(k)


HJ_KNOBS = {
    "hash": {"hj_mode": -1},
    "hash_part": {"hj_mode": 0},
    "hash_part_2pass": {"hj_mode": 0, "hj_onepass": 0},
    "hash_part_rounds": {"hj_mode": 0, "hj_probe": 0},
    "hash_part_rounds_512": {"hj_mode": 0, "hj_probe": 0, "hj_block": 512},
    "hash_global": {"hj_mode": 1},
    "smj_lookback": {"smj_merge": 0},
}


def join_fn(ctx, which):
    """hash = mode picked by build size (hj_mode -1); hash_part = radix-
    partitioned LDS tables (hj_mode 0) with the direct-slot probe (one
    partition pass per side into bucketed tables up to 2^24 build rows);
    hash_part_2pass = the same with the 8-bit radix partition passes and
    linear-probing tables (hj_onepass 0); hash_part_rounds(_512) = the same with the compacting probe in 1024- (512-) thread
    blocks; hash_global = one global table (hj_mode 1); smj = sort-merge join
    (direct-slot merge); smj_lookback = the merge that compacts through a
    decoupled look-back."""
    if which == "smj":
        return ops.sort_merge_join

    def fn(c, *args):
        knobs = HJ_KNOBS[which]
        call = ops.sort_merge_join if which.startswith("smj") else ops.hash_join
        for k, v in knobs.items():
            ctx.set_tune(k, v)
        try:
            return call(c, *args)
        finally:
            for k in knobs:
                ctx.clear_tune(k)
    return fn


@pytest.mark.parametrize("zipf", [False, True])
@pytest.mark.parametrize("which", ["hash", "hash_part", "hash_part_2pass", "hash_part_rounds", "hash_part_rounds_512", "hash_global", "smj",
                                   "smj_lookback"])
def test_join_pk_fk_2_20(ctx, orc, zipf, which):
    rk, rp, sk, sp = orc.join_workload(1 << 17, 1 << 20, zipf=zipf)
    gp, gb = run(ctx, join_fn(ctx, which), rk, rp, sk, sp)
    ep, eb = orc.sort_merge_join(rk, rp, sk, sp) if which.startswith("smj") else orc.hash_join(rk, rp, sk, sp)
    assert len(gp) == len(sk)
    if which.startswith("smj"):   # fully ordered output: bit-exact including order
        assert np.array_equal(gp, ep) and np.array_equal(gb, eb)
    else:
        assert same_multiset((gp, gb), (ep, eb))


@pytest.mark.parametrize("which", ["hash_part", "hash_part_2pass", "hash_part_rounds", "hash_part_rounds_512", "hash_global", "smj",
                                   "smj_lookback"])
@pytest.mark.parametrize("nb,ns,kr,ks", [(1, 1, 1, 1), (3000, 5000, 1000, 2000), (100, 70_000, 60, 200),
                                         (50_000, 3, 10, 10), (4097, 8193, 1 << 31, 1 << 31)])
def test_join_duplicates_misses_edges(ctx, orc, which, nb, ns, kr, ks):
    rk = orc.gen_range(np.uint32, nb + ns, 1, nb, 0, kr - 1)
    rp = orc.gen_u32(nb, 2, nb)
    sk = orc.gen_range(np.uint32, nb + ns, 3, ns, 0, ks - 1)
    sp = orc.gen_u32(ns, 4, ns)
    gp, gb = run(ctx, join_fn(ctx, which), rk, rp, sk, sp)
    ep, eb = orc.sort_merge_join(rk, rp, sk, sp)
    if which.startswith("smj"):
        assert np.array_equal(gp, ep) and np.array_equal(gb, eb)
    else:
        assert same_multiset((gp, gb), (ep, eb))


@pytest.mark.parametrize("which", ["hash_part", "hash_part_2pass", "hash_part_rounds", "hash_global", "smj", "smj_lookback"])
def test_join_unique_build_half_misses(ctx, orc, which):
    """Unique build keys (a permutation of [0, 2^18)) probed by keys uniform
    on [0, 2^19): about half the probe rows match nothing, so the direct
    probe's output has holes that the bitmap packing closes."""
    nb, ns = 1 << 18, 1 << 20
    rk = orc.gen_perm(11, 1, nb, nb)
    rp = orc.gen_u32(nb, 2, nb)
    sk = orc.gen_range(np.uint32, 12, 3, ns, 0, 2 * nb - 1)
    sp = orc.gen_u32(ns, 4, ns)
    gp, gb = run(ctx, join_fn(ctx, which), rk, rp, sk, sp)
    ep, eb = orc.sort_merge_join(rk, rp, sk, sp)
    assert 0.4 * ns < len(ep) < 0.6 * ns
    if which.startswith("smj"):
        assert np.array_equal(gp, ep) and np.array_equal(gb, eb)
    else:
        assert same_multiset((gp, gb), (ep, eb))


def test_join_full_keys_and_max_value(ctx, orc):
    rk = np.array([0, 0xFFFFFFFF, 7, 0x80000000], np.uint32)
    rp = np.arange(4, dtype=np.uint32)
    sk = np.array([0xFFFFFFFF, 0, 0, 5, 0x80000000], np.uint32)
    sp = np.arange(10, 15, dtype=np.uint32)
    for which in ("hash_part", "hash_global", "smj"):
        gp, gb = run(ctx, join_fn(ctx, which), rk, rp, sk, sp)
        assert same_multiset((gp, gb), (np.array([10, 11, 12, 14], np.uint32), np.array([1, 0, 0, 3], np.uint32)))


_FULL = {}


def _full_size_case(ctx, zipf):
    """2^24 build rows, 2^27 probe rows on the device, and the expected output
    multiset from the inverse permutation (no hashing): every probe row
    matches exactly one build row, so the output is exactly
    {(probe payload[j], build payload[pos(probe key[j])])}, held as sorted
    (probe payload << 32 | build payload) words.  Cached per zipf flag."""
    if zipf not in _FULL:
        for (d, _) in _FULL.values():
            for a in d:
                a.free()
        _FULL.clear()
        nr, ns = 1 << 24, 1 << 27
        d = datagen.join_workload(ctx, nr, ns, zipf=zipf)
        hk, hp, sks, sps = (a.numpy() for a in d)
        inv = np.empty(nr, np.int64)
        inv[hk] = np.arange(nr)
        exp = np.sort((sps.astype(np.uint64) << np.uint64(32)) | hp[inv[sks]].astype(np.uint64))
        _FULL[zipf] = (d, exp)
    return _FULL[zipf]


@pytest.mark.parametrize("which", ["hash_part", "hash_part_2pass", "hash_global", "smj"])
@pytest.mark.parametrize("zipf", [False, True])
def test_join_2p24_build_2p27_probe_properties(ctx, zipf, which):
    (rk, rp, sk, sp), exp = _full_size_case(ctx, zipf)
    op, ob, k = join_fn(ctx, which)(ctx, rk, rp, sk, sp)
    assert k == exp.size
    got = np.sort((op.numpy(k).astype(np.uint64) << np.uint64(32)) | ob.numpy(k).astype(np.uint64))
    op.free(); ob.free()
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("which", ["hash", "smj", "smj_lookback"])
def test_join_probe_tile_spans_wide_build_range(ctx, orc, which):
    """Unique build keys, few probe rows spread over all of them: one sorted
    probe tile's build range exceeds the merge's LDS key cache."""
    nb, ns = 100_000, 3_000
    rk = orc.gen_perm(31, 3, nb, nb)
    rp = orc.gen_u32(31, 5, nb)
    sk = orc.gen_range(np.uint32, 31, 4, ns, 0, nb + 500)   # some probe keys miss
    sp = orc.gen_u32(31, 6, ns)
    gp, gb = run(ctx, join_fn(ctx, which), rk, rp, sk, sp)
    ep, eb = orc.sort_merge_join(rk, rp, sk, sp)
    if which.startswith("smj"):
        assert np.array_equal(gp, ep) and np.array_equal(gb, eb)
    else:
        assert same_multiset((gp, gb), (ep, eb))


def _heavy_dup_case(orc):
    """2^17 unique build keys (100 .. 2^17 + 99) plus key 7 repeated 10000
    times and key 13 repeated 9000 times: each heavy key's partition holds
    more build rows than an LDS table (7168) at every partition count.
    Probe: 2^19 keys over the unique range, 20 rows of key 7, 15 of key 13
    and 64 misses."""
    nu = 1 << 17
    rk = np.concatenate([orc.gen_perm(41, 3, nu, nu) + np.uint32(100), np.full(10_000, 7, np.uint32),
                         np.full(9_000, 13, np.uint32)])
    rk = rk[np.argsort(orc.gen_u32(41, 9, rk.size), kind="stable")]
    rp = orc.gen_u32(41, 5, rk.size)
    sk = orc.gen_range(np.uint32, 41, 4, 1 << 19, 100, nu + 99)
    pos = orc.gen_perm(41, 8, 1 << 19, 99)   # 99 distinct probe rows
    sk[pos[:20]] = 7
    sk[pos[20:35]] = 13
    sk[pos[35:]] = np.uint32(1 << 30)
    sp = orc.gen_u32(41, 6, sk.size)
    return rk, rp, sk, sp


@pytest.mark.parametrize("which", ["hash", "hash_part_rounds", "hash_global", "smj"])
def test_join_heavy_duplicate_build_keys_beyond_lds_table(ctx, orc, which):
    """Build partitions larger than an LDS table are joined through a global
    table instead of being refused; every pair is emitted."""
    rk, rp, sk, sp = _heavy_dup_case(orc)
    gp, gb = run(ctx, join_fn(ctx, which), rk, rp, sk, sp)
    ep, eb = orc.sort_merge_join(rk, rp, sk, sp) if which == "smj" else orc.hash_join(rk, rp, sk, sp)
    n_exp = (1 << 19) - 99 + 20 * 10_000 + 15 * 9_000
    assert len(ep) == n_exp
    if which == "smj":
        assert np.array_equal(gp, ep) and np.array_equal(gb, eb)
    else:
        assert same_multiset((gp, gb), (ep, eb))


def test_join_heavy_duplicates_capacity_short(ctx, orc):
    """Device form with a buffer shorter than the output: the full count comes
    back with PIMDAL_E_CAPACITY and the pairs written are a subset of the
    join."""
    import ctypes as C
    from pimdal_amd import _lib
    rk, rp, sk, sp = _heavy_dup_case(orc)
    d = [ctx.to_device(a) for a in (rk, rp, sk, sp)]
    cap = 600_000
    op, ob = ctx.empty(cap, np.uint32), ctx.empty(cap, np.uint32)
    cnt = C.c_uint64()
    rc = _lib.load().pimdal_hash_join_dev(ctx.handle, d[0].vp, d[1].vp, rk.size, d[2].vp, d[3].vp, sk.size, op.vp,
                                          ob.vp, cap, C.byref(cnt))
    ep, eb = orc.hash_join(rk, rp, sk, sp)
    assert rc == -5 and cnt.value == len(ep)
    exp = set(((ep.astype(np.uint64) << np.uint64(32)) | eb).tolist())
    got = (op.numpy().astype(np.uint64) << np.uint64(32)) | ob.numpy()
    assert set(got.tolist()) <= exp


def test_join_tuned_fanout_too_small_doubles(ctx, orc):
    """A tuned partition count far too small for the build side (2 partitions
    of ~2^15 rows): the join doubles it until the partitions fit."""
    rk, rp, sk, sp = orc.join_workload(1 << 16, 1 << 18)
    ctx.set_tune("hj_mode", 0)
    ctx.set_tune("hj_fanout", 2)
    try:
        gp, gb = run(ctx, ops.hash_join, rk, rp, sk, sp)
    finally:
        ctx.clear_tune("hj_mode")
        ctx.clear_tune("hj_fanout")
    ep, eb = orc.hash_join(rk, rp, sk, sp)
    assert same_multiset((gp, gb), (ep, eb))


def test_join_unique_build_above_2p26_rows(ctx):
    """2^26 + 2^20 unique build keys (more than 16384 partitions of 4096 rows)
    and 2^25 probe rows: joined, not refused; the output multiset is the
    closed form from the inverse permutation."""
    nr, ns = (1 << 26) + (1 << 20), 1 << 25
    rk, rp, sk, sp = datagen.join_workload(ctx, nr, ns)
    hk, hp, sks, sps = (a.numpy() for a in (rk, rp, sk, sp))
    op, ob, k = ops.hash_join(ctx, rk, rp, sk, sp)
    assert k == ns
    inv = np.empty(nr, np.int64)
    inv[hk] = np.arange(nr)
    exp = np.sort((sps.astype(np.uint64) << np.uint64(32)) | hp[inv[sks]].astype(np.uint64))
    got = np.sort((op.numpy(k).astype(np.uint64) << np.uint64(32)) | ob.numpy(k).astype(np.uint64))
    for a in (rk, rp, sk, sp, op, ob):
        a.free()
    assert np.array_equal(got, exp)


def _ref_hash_low(keys, bits):
    """ref_hash(key) & (2^bits - 1): low 32 bits of the byte-swapped 64-bit
    product with 0x9E3779B185EBCA87 (the partition number)."""
    p = keys.astype(np.uint64) * np.uint64(11400714785074694791)
    return (p.byteswap() & np.uint64(0xFFFFFFFF)).astype(np.uint64) & np.uint64((1 << bits) - 1)


def test_join_one_pass_partition_past_bucketed_table(ctx, orc):
    """2^20 unique build keys (64 one-pass partitions) of which 17000 hash to
    partition 5, past the 16896 rows a bucketed table holds: the join falls
    through to the general path and still matches the oracle."""
    cand = np.arange(1, 1 << 24, dtype=np.uint32)
    hot = cand[_ref_hash_low(cand, 6) == 5][:17000]
    rest = cand[_ref_hash_low(cand, 6) != 5][: (1 << 20) - hot.size]
    rk = np.concatenate([hot, rest])
    rk = rk[np.random.default_rng(3).permutation(rk.size)]
    rp = np.arange(rk.size, dtype=np.uint32)
    sk = rk[np.random.default_rng(4).integers(0, rk.size, 1 << 22)]
    sp = np.arange(sk.size, dtype=np.uint32)
    exp = orc.hash_join(rk, rp, sk, sp)
    got = run(ctx, join_fn(ctx, "hash_part"), rk, rp, sk, sp)
    assert same_multiset(got, exp)


@pytest.mark.parametrize("nb", [1, 2, 16384, 16385, (1 << 24) + 1])
def test_join_one_pass_fanout_edges(ctx, orc, nb):
    """Build sizes at the one-pass partition counts' edges (2 partitions,
    2 -> 4, and past 1024 partitions, where the radix passes take over)."""
    rk = orc.gen_perm(71, 3, nb, nb)
    rp = orc.gen_u32(71, 5, nb)
    ns = 3 * nb + 7 if nb < (1 << 20) else 1 << 22
    sk = orc.gen_range(np.uint32, 71, 4, ns, 0, nb + nb // 3)   # a quarter of the probe rows miss
    sp = orc.gen_u32(71, 6, ns)
    exp = orc.hash_join(rk, rp, sk, sp)
    got = run(ctx, join_fn(ctx, "hash_part"), rk, rp, sk, sp)
    assert same_multiset(got, exp)
