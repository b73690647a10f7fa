"""SQL null semantics of the column-level operators on the GPU: NULL keys
select and join nothing, payload validity follows the rows, NULL keys form
one group, SUM skips NULL values and COUNT counts non-NULL ones.  Expected
results come from the oracle run on the non-NULL rows; on tiny inputs the
semantics are pinned by pure-Python loops."""
import numpy as np
import pyarrow as pa
import pytest

from pimdal_amd import ingest

pytestmark = pytest.mark.gpu


def _nullable(values, valid):
    return pa.array(values, mask=~np.asarray(valid, dtype=bool))


def _imp(ctx, arr):
    return ingest.import_arrow(ctx, arr)


def _list(ctx, col):
    return ingest.column_to_arrow(ctx, col).to_pylist() if col is not None else None


def _valid(n, seed, p_null):
    return np.random.default_rng(seed).random(n) >= p_null


# ---------------- select ----------------

def _select_expected(orc, k, kv, p, pv, lo, hi):
    pay, rid = orc.select(k, p, lo, hi)
    keep = kv[rid]
    rid = rid[keep]
    return [int(p[r]) if pv[r] else None for r in rid], rid.tolist()


@pytest.mark.parametrize("n,p_null", [(1, 0.5), (1000, 0.3), (300_001, 0.1), (300_001, 1.0)])
def test_select_with_nulls(ctx, orc, n, p_null):
    k, p = orc.select_workload(n, seed=61 + n)
    kv, pv = _valid(n, 1, p_null), _valid(n, 2, 0.2)
    kc, pc = _imp(ctx, _nullable(k, kv)), _imp(ctx, _nullable(p, pv))
    op, orow = ingest.col_select(ctx, kc, pc, 0, 1 << 31)
    exp_p, exp_r = _select_expected(orc, k, kv, p, pv, 0, 1 << 31)
    assert _list(ctx, orow) == exp_r and _list(ctx, op) == exp_p


def test_select_with_nulls_tiny_brute_force(ctx):
    k = np.array([5, 15, 25, 35, 12, 18], np.int32)
    kv = [True, False, True, True, True, True]
    p = np.arange(100, 106, dtype=np.uint32)
    pv = [True, True, True, False, True, False]
    op, orow = ingest.col_select(ctx, _imp(ctx, _nullable(k, kv)), _imp(ctx, _nullable(p, pv)), 10, 30)
    rows = [i for i in range(6) if kv[i] and 10 <= k[i] < 30]
    assert _list(ctx, orow) == rows
    assert _list(ctx, op) == [int(p[i]) if pv[i] else None for i in rows]


# ---------------- joins ----------------

@pytest.mark.parametrize("sort_merge", [False, True], ids=["hash", "smj"])
@pytest.mark.parametrize("nb,ns", [(50, 200), (100_000, 400_000)])
def test_join_with_nulls(ctx, orc, sort_merge, nb, ns):
    rk, rp, sk, sp = orc.join_workload(nb, ns)
    rkv, skv, rpv, spv = _valid(nb, 3, 0.2), _valid(ns, 4, 0.25), _valid(nb, 5, 0.1), _valid(ns, 6, 0.1)
    op, ob = ingest.col_join(ctx, _imp(ctx, _nullable(rk, rkv)), _imp(ctx, _nullable(rp, rpv)),
                             _imp(ctx, _nullable(sk, skv)), _imp(ctx, _nullable(sp, spv)), sort_merge)
    brow = np.nonzero(rkv)[0].astype(np.uint32)
    prow = np.nonzero(skv)[0].astype(np.uint32)
    jp, jb = orc.hash_join(rk[brow], brow, sk[prow], prow)       # (probe row, build row) of every pair
    exp = sorted(zip([int(sp[r]) if spv[r] else None for r in jp], [int(rp[r]) if rpv[r] else None for r in jb]),
                 key=lambda t: (t[0] is None, t[0] or 0, t[1] is None, t[1] or 0))
    got = sorted(zip(_list(ctx, op), _list(ctx, ob)), key=lambda t: (t[0] is None, t[0] or 0, t[1] is None, t[1] or 0))
    assert got == exp and len(got) == int(skv[np.isin(sk, rk[rkv])].sum())


def test_join_with_null_keys_tiny_brute_force(ctx):
    bk, bkv = np.array([1, 2, 2, 3], np.uint32), [True, True, False, True]
    pk, pkv = np.array([2, 3, 1, 1, 2], np.uint32), [True, False, True, True, True]
    bp, pp = np.array([10, 20, 21, 30], np.uint32), np.array([100, 101, 102, 103, 104], np.uint32)
    op, ob = ingest.col_join(ctx, _imp(ctx, _nullable(bk, bkv)), _imp(ctx, pa.array(bp)),
                             _imp(ctx, _nullable(pk, pkv)), _imp(ctx, pa.array(pp)))
    exp = sorted((int(pp[j]), int(bp[i])) for i in range(4) for j in range(5)
                 if bkv[i] and pkv[j] and bk[i] == pk[j])
    assert sorted(zip(_list(ctx, op), _list(ctx, ob))) == exp


# ---------------- aggregation ----------------

def _aggr_expected(orc, k, kv, v, vv):
    keys = k[kv]
    v0 = np.where(vv, v, 0)[kv]
    nn = vv.astype(np.int64)[kv]
    gk, gs, _ = orc.sort_aggr(keys, v0)
    _, gn, _ = orc.sort_aggr(keys, nn)
    rows = [(int(a), int(s) if c else None, int(c)) for a, s, c in zip(gk, gs, gn)]
    if (~kv).any():
        m = ~kv
        c = int(vv[m].sum())
        s = int(np.sum(np.where(vv, v, 0)[m], dtype=np.int64))
        rows.append((None, s if c else None, c))
    return rows


@pytest.mark.parametrize("sort", [False, True], ids=["hash", "sort"])
@pytest.mark.parametrize("n,groups", [(1000, 30), (500_000, 20_000)])
def test_aggregation_with_nulls(ctx, orc, sort, n, groups):
    k = orc.gen_range(np.uint32, 71, 1, n, 0, groups - 1)
    v = orc.gen_range(np.int64, 71, 2, n, -(1 << 40), 1 << 40)
    kv, vv = _valid(n, 7, 0.05), _valid(n, 8, 0.3)
    vv[k == 3] = False                                   # a group whose values are all NULL
    ok, os_, oc = ingest.col_aggr(ctx, _imp(ctx, _nullable(k, kv)), _imp(ctx, _nullable(v, vv)), sort)
    got = list(zip(_list(ctx, ok), _list(ctx, os_), _list(ctx, oc)))
    assert got == _aggr_expected(orc, k, kv, v, vv)


def test_aggregation_with_nulls_tiny_brute_force(ctx):
    k = np.array([1, 2, 1, 9, 2, 3, 1], np.uint32)
    kv = [True, True, True, False, True, True, False]
    v = np.array([5, 7, -2, 4, 8, 6, 1], np.int64)
    vv = [True, False, True, True, True, False, False]
    ok, os_, oc = ingest.col_aggr(ctx, _imp(ctx, _nullable(k, kv)), _imp(ctx, _nullable(v, vv)))
    groups = {}
    for i in range(7):
        g = int(k[i]) if kv[i] else None
        s, c = groups.get(g, (0, 0))
        groups[g] = (s + (int(v[i]) if vv[i] else 0), c + (1 if vv[i] else 0))
    exp = [(g, s if c else None, c) for g, (s, c) in sorted(groups.items(), key=lambda t: (t[0] is None, t[0] or 0))]
    assert list(zip(_list(ctx, ok), _list(ctx, os_), _list(ctx, oc))) == exp


@pytest.mark.parametrize("n,p_null", [(1, 0.5), (4097, 0.3), (300_001, 0.1)])
def test_select_uint8_key_column_with_nulls(ctx, orc, n, p_null):
    """A nullable uint8 key column (an Arrow uint8 array, such as a flag
    column) through the column select: the oracle's uint32 selection over
    the same values, NULL keys kept out, payload validity following the
    rows."""
    rng = np.random.default_rng(90 + n)
    k = rng.integers(0, 256, n, dtype=np.uint16).astype(np.uint8)
    p = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    kv, pv = _valid(n, 7, p_null), _valid(n, 8, 0.2)
    kc, pc = _imp(ctx, _nullable(k, kv)), _imp(ctx, _nullable(p, pv))
    op, orow = ingest.col_select(ctx, kc, pc, 3, 200)
    exp_p, exp_r = _select_expected(orc, k.astype(np.uint32), kv, p, pv, 3, 200)
    assert _list(ctx, orow) == exp_r and _list(ctx, op) == exp_p
