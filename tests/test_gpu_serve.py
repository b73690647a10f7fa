"""Serving loop: continuous batching vs one query at a time, results equal to
each other, to each query run alone, and to the host restatement."""
import numpy as np
import pytest

from pimdal_amd import datagen, serving, PimdalError
from tests.serve_check import expected, expected_oracle

pytestmark = pytest.mark.gpu

N_FACT, N_DIM = 1 << 22, 1 << 20


@pytest.fixture(params=[1, 0], ids=["join_index", "probe_index"], autouse=True)
def join_mode(request, ctx):
    """Batched joins read the resident foreign-key join index (1) or probe
    the dimension-key index per row (0)."""
    ctx.set_tune("serve_join_index", request.param)
    yield request.param
    ctx.set_tune("serve_join_index", 1)


@pytest.fixture(scope="module")
def tables(ctx):
    fact = datagen.fact_table(ctx, N_FACT, groups=N_DIM)
    dim = datagen.dim_table(ctx, N_DIM)
    host_fact = {c: v.numpy() for c, v in fact.items()}
    host_dim = {c: v.numpy() for c, v in dim.items()}
    order = np.argsort(host_dim["key"], kind="stable")
    return fact, dim, host_fact, host_dim, (order, host_dim["key"][order])


def test_batched_equals_sequential_and_host(ctx, orc, tables):
    fact, dim, hf, hd, ds = tables
    srv = serving.Server(ctx, fact, dim, slots=4, heap_bytes=256 << 20)
    q = serving.gen_queries(400, N_FACT, rate_qps=2e5, min_log2=8, max_log2=20, seed=7)
    r_seq, st_seq = srv.run(q, serving.SEQUENTIAL)
    r_bat, st_bat = srv.run(q, serving.BATCHED)
    assert np.array_equal(r_seq["count"], r_bat["count"])
    assert np.array_equal(r_seq["checksum"], r_bat["checksum"])
    assert st_bat["batches"] < len(q)          # small queries were coalesced
    for i in range(len(q)):
        cnt, ck = expected(q[i], hf, hd, orc, ds)
        assert (int(r_bat["count"][i]), int(r_bat["checksum"][i])) == (cnt, ck), (i, q[i])
    srv.close()


def test_each_result_same_alone_and_in_the_mix(ctx, tables):
    fact, dim, hf, hd, ds = tables
    srv = serving.Server(ctx, fact, dim, slots=3, heap_bytes=256 << 20)
    q = serving.gen_queries(120, N_FACT, rate_qps=0, min_log2=12, max_log2=21, seed=11)  # all arrive at once
    mix, st = srv.run(q, serving.BATCHED)
    assert st["max_inflight"] > 1
    for i in range(0, len(q), 7):
        alone, _ = srv.run(q[i:i + 1], serving.BATCHED)
        assert (alone["count"][0], alone["checksum"][0]) == (mix["count"][i], mix["checksum"][i])
    srv.close()


def test_admission_rejects_query_larger_than_the_heap(ctx, tables):
    fact, dim, *_ = tables
    srv = serving.Server(ctx, fact, dim, slots=1, heap_bytes=1 << 20)
    q = serving.gen_queries(1, N_FACT, rate_qps=0, min_log2=20, max_log2=20)
    q["kind"] = serving.JOIN
    with pytest.raises(PimdalError, match="larger than the serving heap"):
        srv.run(q, serving.BATCHED)
    srv.close()


def test_admission_by_footprint_in_a_small_heap(ctx, orc, tables):
    """A heap that holds only a few of the queries at once: batches are
    admitted as their footprint frees up, the heap's peak use stays within
    it, and every result equals the host restatement."""
    fact, dim, hf, hd, ds = tables
    heap = 24 << 20
    srv = serving.Server(ctx, fact, dim, slots=4, heap_bytes=heap)
    q = serving.gen_queries(80, N_FACT, rate_qps=0, min_log2=18, max_log2=21, seed=41)   # up to 16 MiB each
    r, st = srv.run(q, serving.BATCHED)
    assert st["heap_bytes"] == heap and 0 < st["heap_peak_bytes"] <= heap
    assert st["batches"] > 4
    for i in range(len(q)):
        assert (int(r["count"][i]), int(r["checksum"][i])) == expected(q[i], hf, hd, orc, ds), i
    srv.close()


def test_tenants_take_turns(ctx, orc, tables):
    """Tenant 0 queues 24 large queries before tenant 1's 24 small ones, all
    at once, into a heap that holds about two large ones: tenants are
    admitted in turn, so tenant 1 does not wait behind all of tenant 0, and
    every result is right."""
    fact, dim, hf, hd, ds = tables
    big = serving.gen_queries(24, N_FACT, rate_qps=0, min_log2=21, max_log2=21, seed=51)
    small = serving.gen_queries(24, N_FACT, rate_qps=0, min_log2=12, max_log2=12, seed=52)
    big["kind"] = serving.JOIN
    small["tenant"] = 1
    q = np.concatenate([big, small])
    srv = serving.Server(ctx, fact, dim, slots=2, heap_bytes=40 << 20)
    r, st = srv.run(q, serving.BATCHED)
    assert st["tenants"] == 2
    t_big = np.sort(r["end_us"][:24])
    t_small = np.sort(r["end_us"][24:])
    assert t_small[11] < t_big[-1]    # half of tenant 1 done before tenant 0 finishes
    for i in range(len(q)):
        assert (int(r["count"][i]), int(r["checksum"][i])) == expected(q[i], hf, hd, orc, ds), i
    srv.close()


def test_dimension_keys_must_be_unique(ctx, tables):
    fact, dim, hf, hd, _ = tables
    k = hd["key"].copy()
    k[5] = k[6]
    dup = dict(dim)
    dup["key"] = ctx.to_device(k)
    with pytest.raises(PimdalError, match="unique"):
        serving.Server(ctx, fact, dup, slots=1, heap_bytes=16 << 20)


def test_join_probes_that_miss(ctx, orc, tables):
    """Half the dimension keys replaced by values no fact row holds: those
    fact rows find no dimension row (the probe walks to a bucket that never
    overflowed), in both modes and in the host restatement."""
    fact, dim, hf, hd, _ = tables
    k = hd["key"].copy()
    k[::2] = -1 - np.arange(k[::2].size, dtype=np.int32)
    d2 = dict(dim)
    d2["key"] = ctx.to_device(k)
    hd2 = dict(hd)
    hd2["key"] = k
    order = np.argsort(k, kind="stable")
    ds2 = (order, k[order])
    srv = serving.Server(ctx, fact, d2, slots=2, heap_bytes=256 << 20)
    q = serving.gen_queries(60, N_FACT, rate_qps=0, min_log2=10, max_log2=18, seed=23)
    q["kind"] = serving.JOIN
    r_seq, _ = srv.run(q, serving.SEQUENTIAL)
    r_bat, _ = srv.run(q, serving.BATCHED)
    assert np.array_equal(r_seq["count"], r_bat["count"]) and np.array_equal(r_seq["checksum"], r_bat["checksum"])
    for i in range(len(q)):
        assert (int(r_bat["count"][i]), int(r_bat["checksum"][i])) == expected(q[i], hf, hd2, orc, ds2), i
    srv.close()


def test_serving_full_size_against_the_oracle(ctx, orc):
    """Config 5 at its full size: 6e7-row fact table, 1.5e7-row dimension
    table, slices up to 2^26 rows.  A sample of the stream's queries (each
    kind) is checked against the CPU restatement, select row ids in output
    order, join pairs in any order, group sums; the whole stream's batched
    results equal the one-at-a-time ones."""
    n_fact, n_dim = 60_000_000, 15_000_000
    fact = datagen.fact_table(ctx, n_fact, groups=n_dim)
    dim = datagen.dim_table(ctx, n_dim)
    hf = {c: v.numpy() for c, v in fact.items()}
    hd = {c: v.numpy() for c, v in dim.items()}
    srv = serving.Server(ctx, fact, dim, slots=4, heap_bytes=4 << 30)
    q = serving.gen_queries(600, n_fact, rate_qps=5e3, seed=0x5EED0005)
    r_bat, st = srv.run(q, serving.BATCHED)
    r_seq, _ = srv.run(q, serving.SEQUENTIAL)
    assert np.array_equal(r_seq["count"], r_bat["count"]) and np.array_equal(r_seq["checksum"], r_bat["checksum"])
    sample = []
    for kind in (serving.SELECT, serving.JOIN, serving.GROUPBY):
        idx = np.nonzero(q["kind"] == kind)[0]
        big = idx[np.argsort(q["rows"][idx])[::-1][:2]]          # the largest slices of each kind
        sample += list(big) + list(idx[:3])
    assert q["rows"].max() > (1 << 25)
    for i in sample:
        assert (int(r_bat["count"][i]), int(r_bat["checksum"][i])) == expected_oracle(q[i], hf, hd, orc), (i, q[i])
    srv.close()
    for d in (fact, dim):
        for a in d.values():
            a.free()
