"""The CPU restatement pinned against the reference's own CPU operators
(reference/db_operators/*.cpp built into oracle/_ref by oracle/ref.mk) and
the PIM library's hash mixers (pimdal/hash/hash_func.h, same recipe).

Inputs avoid the reference's documented quirks: the row count is a multiple
of its 8000-row block and equals its compiled BUFFER_SIZE (select.cpp:15-16
drops a ragged tail), the last select key satisfies the predicate (select.cpp:
69-73 otherwise writes one element past the output), the select runs on one
thread (the same lines race at block boundaries between threads), and every
hash-join probe key matches (hash_join.cpp:148-153 places output at the
prefix of probe-partition sizes).  Skipped where oracle/_ref was not built."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from tests.util import canon

REF = Path(__file__).resolve().parents[1] / "oracle" / "_ref"
N = 800_000


def ref(name):
    p = REF / f"ref_{name}.so"
    if not p.exists():
        pytest.skip("oracle/_ref not built (needs the reference tree and pyarrow's Arrow C++)")
    return C.CDLL(str(p))


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def test_select_matches_reference(orc):
    L = ref("select")
    L.ref_select_i32.restype = C.c_uint64
    L.ref_select_i32.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]
    k = orc.gen_range(np.int32, 11, 1, N, 1, 50)   # reference data shape (select.cpp:18-31)
    k[-1] = 15
    out = np.empty(N + 16, np.int32)
    nt = orc.num_threads()
    orc.set_threads(1)
    try:
        cnt = L.ref_select_i32(ptr(k), N, ptr(out), N + 16)
    finally:
        orc.set_threads(nt)
    op, orow = orc.select(k, k.view(np.uint32), 10, 30)
    assert cnt == len(op)
    assert np.array_equal(out[:cnt], op.view(np.int32))


def test_sort_matches_reference(orc):
    L = ref("sort")
    L.ref_sort_u32.restype = C.c_uint64
    L.ref_sort_u32.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
    k = orc.gen_u32(12, 1, N)
    out = np.empty(N, np.uint32)
    assert L.ref_sort_u32(ptr(k), N, ptr(out)) == N
    assert np.array_equal(out, orc.sort_keys(k))
    ok, ov = orc.sort_pairs(k, np.arange(N, dtype=np.uint32))
    assert np.array_equal(out, ok)


def test_hash_join_matches_reference_in_order(orc):
    L = ref("hash_join")
    L.ref_hash_join_u32.restype = C.c_uint64
    L.ref_hash_join_u32.argtypes = [C.c_void_p] * 2 + [C.c_uint64] + [C.c_void_p] * 2
    inner = orc.gen_perm(13, 1, N, N) + np.uint32(1)             # shuffled 1..N (hash_join.cpp:27-40)
    outer = orc.gen_range(np.uint32, 13, 2, N, 1, N)             # uniform [1, N] (hash_join.cpp:42-55)
    oi = np.empty(N, np.uint32)
    oo = np.empty(N, np.uint32)
    assert L.ref_hash_join_u32(ptr(inner), ptr(outer), N, ptr(oi), ptr(oo)) == N
    # restatement with the reference's partition count; payload = key
    op, ob = orc.hash_join(inner, inner, outer, outer, part_n=16384)
    # same partition function and in-bucket order, so the output order agrees too
    assert np.array_equal(oo, op) and np.array_equal(oi, ob)


def test_hash_aggr_matches_reference(orc):
    L = ref("aggregate")
    L.ref_hash_aggr.restype = C.c_uint64
    L.ref_hash_aggr.argtypes = [C.c_void_p] * 2 + [C.c_uint64] + [C.c_void_p] * 2 + [C.c_uint64]
    k = orc.gen_range(np.uint32, 14, 1, N, 1, 50)     # aggregate.cpp:18-37 shape
    v = orc.gen_range(np.int32, 14, 2, N, 1, 10)
    ok = np.empty(64, np.uint32)
    ov = np.empty(64, np.int32)
    g = L.ref_hash_aggr(ptr(k), ptr(v), N, ptr(ok), ptr(ov), 64)
    rk, rv = canon(ok[:g], ov[:g])
    ek, es, ec = orc.hash_aggr(k, v.astype(np.int64))
    assert np.array_equal(rk, ek) and np.array_equal(rv.astype(np.int64), es)


# ---------------- the PIM library's hash mixers (pimdal/hash/hash_func.h) ----------------

def _ref_hash():
    L = ref("hash_func")
    L.ref_pim_hash.restype = C.c_uint32
    L.ref_pim_hash.argtypes = [C.c_int, C.c_uint32]
    return L


def test_pim_hash_mixers_match_reference(orc):
    L = _ref_hash()
    keys = [0, 1, 2, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF] + \
        np.random.default_rng(31).integers(0, 1 << 32, 20000, dtype=np.uint64).tolist()
    for w in range(5):
        for k in keys:
            assert orc.pim_hash(w, k) == L.ref_pim_hash(w, k), (w, k)


def test_bloom_filters_follow_reference_hashes(orc):
    """The oracle's filters against bits placed with the reference's own
    mixers: partition = bits [shift - part_bits, shift) of hash0, bit
    hash_func[f](key) % 16384 of that partition's filter, byte layout of
    insert_bloom_filter (pimdal/hash/hash_steps.c:143-152)."""
    L = _ref_hash()
    part_bits, shift = 3, 29
    keys = np.random.default_rng(32).integers(0, 1 << 32, 3000, dtype=np.uint64).astype(np.uint32)
    exp = np.zeros((1 << part_bits) * 2048, np.uint8)
    for k in keys.tolist():
        p = (L.ref_pim_hash(0, k) >> (shift - part_bits)) & ((1 << part_bits) - 1)
        for f in range(1, 5):
            bit = L.ref_pim_hash(f, k) % 16384
            exp[p * 2048 + bit // 8] |= 1 << (bit % 8)
    assert np.array_equal(orc.bloom_build(keys, part_bits, shift), exp)
