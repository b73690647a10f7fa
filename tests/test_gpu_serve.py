"""Serving loop: continuous batching vs one query at a time, results equal to
each other, to each query run alone, and to the host restatement."""
import numpy as np
import pytest

from pimdal_amd import datagen, serving, PimdalError
from tests.serve_check import expected

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
    srv = serving.Server(ctx, fact, dim, slots=4, arena_bytes=256 << 20)
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
    srv = serving.Server(ctx, fact, dim, slots=3, arena_bytes=256 << 20)
    q = serving.gen_queries(120, N_FACT, rate_qps=0, min_log2=12, max_log2=21, seed=11)  # all arrive at once
    mix, st = srv.run(q, serving.BATCHED)
    assert st["max_inflight"] > 1
    for i in range(0, len(q), 7):
        alone, _ = srv.run(q[i:i + 1], serving.BATCHED)
        assert (alone["count"][0], alone["checksum"][0]) == (mix["count"][i], mix["checksum"][i])
    srv.close()


def test_admission_rejects_query_larger_than_a_slot(ctx, tables):
    fact, dim, *_ = tables
    srv = serving.Server(ctx, fact, dim, slots=1, arena_bytes=1 << 20)
    q = serving.gen_queries(1, N_FACT, rate_qps=0, min_log2=20, max_log2=20)
    q["kind"] = serving.JOIN
    with pytest.raises(PimdalError, match="larger than a slot"):
        srv.run(q, serving.BATCHED)
    srv.close()


def test_dimension_keys_must_be_unique(ctx, tables):
    fact, dim, hf, hd, _ = tables
    k = hd["key"].copy()
    k[5] = k[6]
    dup = dict(dim)
    dup["key"] = ctx.to_device(k)
    with pytest.raises(PimdalError, match="unique"):
        serving.Server(ctx, fact, dup, slots=1, arena_bytes=16 << 20)


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
    srv = serving.Server(ctx, fact, d2, slots=2, arena_bytes=256 << 20)
    q = serving.gen_queries(60, N_FACT, rate_qps=0, min_log2=10, max_log2=18, seed=23)
    q["kind"] = serving.JOIN
    r_seq, _ = srv.run(q, serving.SEQUENTIAL)
    r_bat, _ = srv.run(q, serving.BATCHED)
    assert np.array_equal(r_seq["count"], r_bat["count"]) and np.array_equal(r_seq["checksum"], r_bat["checksum"])
    for i in range(len(q)):
        assert (int(r_bat["count"][i]), int(r_bat["checksum"][i])) == expected(q[i], hf, hd2, orc, ds2), i
    srv.close()
