"""The CPU restatement (oracle/) pinned against closed forms, brute force and
the mathematics of each operator.  The reference ships no golden vectors
or tests for this path, so each check is chosen so that a plausible slip
(dropped tail, wrong index, unstable tie order, lost duplicate, overflow)
fails it.  Runs on the CPU."""
import numpy as np
import pytest

from tests.util import canon


# ---------------- selection ----------------

@pytest.mark.parametrize("n", [0, 1, 7999, 8000, 8001, 16000, 100003])
def test_select_matches_brute_force(orc, n):
    k, p = orc.select_workload(n, seed=1234 + n)
    lo, hi = 100_000_000, 1_500_000_000
    op, orow = orc.select(k, p, lo, hi)
    idx = [i for i in range(n) if lo <= int(k[i]) < hi] if n < 20000 else np.nonzero((k >= lo) & (k < hi))[0]
    assert list(orow) == list(idx)          # tail rows included, ascending order
    assert np.array_equal(op, p[np.asarray(idx, dtype=np.int64)])


def test_select_reference_predicate_int32(orc):
    # reference/db_operators/select.cpp:33-36 predicate 10 <= x < 30 on int32 keys in [1, 50]
    k = orc.gen_range(np.int32, 7, 1, 50_000, 1, 50)
    op, orow = orc.select(k, k.view(np.uint32), 10, 30)
    ref = k[(k >= 10) & (k < 30)]
    assert np.array_equal(op.view(np.int32), ref)
    assert np.all(np.diff(orow.astype(np.int64)) > 0)


def test_select_2_20_selectivity_10pct(orc):
    k, p = orc.select_workload(1 << 20)
    op, orow = orc.select(k, p, 0, 429496730)
    assert abs(len(op) / (1 << 20) - 0.1) < 0.003
    assert len(op) == int(np.count_nonzero(k < 429496730))


# ---------------- sort ----------------

@pytest.mark.parametrize("n", [0, 1, 2, 1000, 65537])
def test_sort_pairs_is_the_stable_order(orc, n):
    k = orc.gen_range(np.uint32, 99, 1, n, 0, 300)  # many ties
    v = orc.gen_u32(99, 2, n)
    ok, ov = orc.sort_pairs(k, v)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(ok, k[order]) and np.array_equal(ov, v[order])


def test_sort_ties_keep_row_order(orc):
    k, rid = orc.sort_workload(200_000, zipf=True, distinct=1000)
    ok, ov = orc.sort_pairs(k, rid)
    assert np.all(ok[:-1] <= ok[1:])
    same = ok[:-1] == ok[1:]
    assert np.all(ov[:-1][same] < ov[1:][same])
    assert np.array_equal(np.sort(ov), rid)   # a permutation of the rows


def test_sort_keys_only_matches_reference_sort(orc):
    k = orc.gen_u32(5, 1, 100_000)
    assert np.array_equal(orc.sort_keys(k), np.sort(k))
    ok, _ = orc.sort_pairs(k)
    assert np.array_equal(ok, np.sort(k))


def test_sort_u64_payload(orc):
    k = orc.gen_range(np.uint32, 3, 1, 5000, 0, 50)
    v = orc.gen_range(np.int64, 3, 2, 5000, -(1 << 40), 1 << 40).view(np.uint64)
    ok, ov = orc.sort_pairs(k, v)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(ov, v[order])


# ---------------- joins ----------------

def _closed_form_pkfk(rk, rp, sk, sp):
    inv = np.empty(rk.size, np.int64)
    inv[rk] = np.arange(rk.size)
    return sp, rp[inv[sk]]


@pytest.mark.parametrize("zipf", [False, True])
def test_hash_join_pk_fk_closed_form(orc, zipf):
    rk, rp, sk, sp = orc.join_workload(1 << 12, 1 << 15, zipf=zipf, seed=77)
    op, ob = orc.hash_join(rk, rp, sk, sp, part_n=64)
    ep, eb = _closed_form_pkfk(rk, rp, sk, sp)
    assert len(op) == len(sk)
    assert all(np.array_equal(a, b) for a, b in zip(canon(op, ob), canon(ep, eb)))


def _brute_join(rk, rp, sk, sp):
    d = {}
    for k, p in zip(rk.tolist(), rp.tolist()):
        d.setdefault(k, []).append(p)
    o1, o2 = [], []
    for k, p in zip(sk.tolist(), sp.tolist()):
        for q in d.get(k, []):
            o1.append(p)
            o2.append(q)
    return np.array(o1, np.uint32), np.array(o2, np.uint32)


def test_hash_join_duplicates_and_misses(orc):
    rk = orc.gen_range(np.uint32, 11, 1, 3000, 0, 999)   # duplicate build keys
    rp = orc.gen_u32(11, 2, 3000)
    sk = orc.gen_range(np.uint32, 11, 3, 5000, 0, 1999)  # half the probes miss
    sp = orc.gen_u32(11, 4, 5000)
    op, ob = orc.hash_join(rk, rp, sk, sp, part_n=16)
    ep, eb = _brute_join(rk, rp, sk, sp)
    assert all(np.array_equal(a, b) for a, b in zip(canon(op, ob), canon(ep, eb)))


def test_hash_join_partition_hash_is_reference_hash(orc):
    # reference/db_operators/hash_join.cpp:20-25 on hand-computed values
    C = 11400714785074694791
    for x in (0, 1, 2, 12345, 2**31, 2**32 - 1):
        p = (x * C) % 2**64
        assert orc.hash32(x) == int.from_bytes(p.to_bytes(8, "little")[::-1], "little") % 2**32


def test_sort_merge_join_equals_hash_join_and_is_key_ordered(orc):
    rk = orc.gen_range(np.uint32, 21, 1, 4000, 0, 2999)
    rp = orc.gen_u32(21, 2, 4000)
    sk = orc.gen_range(np.uint32, 21, 3, 9000, 0, 3999)
    sp = orc.gen_u32(21, 4, 9000)
    mp, mb = orc.sort_merge_join(rk, rp, sk, sp)
    hp, hb = orc.hash_join(rk, rp, sk, sp, part_n=32)
    assert all(np.array_equal(a, b) for a, b in zip(canon(mp, mb), canon(hp, hb)))
    # SMJ emits in (key, probe row, build row) order
    ep, eb = _brute_join(rk, rp, sk, sp)
    order_s = np.argsort(sk, kind="stable")
    exp_p, exp_b = [], []
    d = {}
    ro = np.argsort(rk, kind="stable")
    for i in ro:
        d.setdefault(int(rk[i]), []).append(int(rp[i]))
    for j in order_s:
        for q in d.get(int(sk[j]), []):
            exp_p.append(int(sp[j]))
            exp_b.append(q)
    assert mp.tolist() == exp_p and mb.tolist() == exp_b


# ---------------- aggregation ----------------

def test_hash_aggr_matches_bincount(orc):
    k = orc.gen_range(np.uint32, 31, 1, 200_000, 1, 50)       # reference/db_operators/aggregate.cpp:33-36 shape
    v = orc.gen_range(np.int64, 31, 2, 200_000, 1, 10)
    ok, osum, ocnt = orc.hash_aggr(k, v)
    assert np.array_equal(ok, np.arange(1, 51, dtype=np.uint32))
    assert np.array_equal(osum, np.bincount(k, weights=v, minlength=51)[1:].astype(np.int64))
    assert np.array_equal(ocnt, np.bincount(k, minlength=51)[1:])
    sk, ss, sc = orc.sort_aggr(k, v)
    assert np.array_equal(sk, ok) and np.array_equal(ss, osum) and np.array_equal(sc, ocnt)


def test_hash_aggr_high_cardinality_and_negative(orc):
    k = orc.gen_u32(41, 1, 50_000) % np.uint32(20011)
    v = orc.gen_range(np.int64, 41, 2, 50_000, -(1 << 45), 1 << 45)
    ok, osum, ocnt = orc.hash_aggr(k, v)
    u, inv = np.unique(k, return_inverse=True)
    s = np.zeros(u.size, dtype=object)
    for i, x in zip(inv.tolist(), v.tolist()):
        s[i] += x
    assert np.array_equal(ok, u) and [int(x) for x in osum] == list(s)
    assert np.array_equal(ocnt, np.bincount(inv))


def _q1_brute(t, cutoff, by_key):
    groups = {}
    for i in range(t["date"].size):
        if t["date"][i] > cutoff:
            continue
        g = int(t["orderkey"][i]) if by_key else int(t["flag1"][i]) * 256 + int(t["flag2"][i])
        a = groups.setdefault(g, [0, 0, 0, 0, 0, 0])
        q, p, d, x = (int(t[c][i]) for c in ("quantity", "price", "discount", "tax"))
        dp = p * (100 - d)
        a[0] += q; a[1] += p; a[2] += d; a[3] += dp; a[4] += dp * (100 + x); a[5] += 1
    return groups


@pytest.mark.parametrize("by_key", [False, True])
def test_q1_matches_python_integers(orc, by_key):
    t = orc.fact_table(20_000, seed=123, groups=3000)
    cutoff = 2505
    out = orc.q1_key(t, cutoff) if by_key else orc.q1_flags(t, cutoff)
    exp = _q1_brute(t, cutoff, by_key)
    keys = sorted(exp)
    assert list(out["key"]) == keys
    for j, g in enumerate(keys):
        e = exp[g]
        got = [int(out["sum_qty"][j]), int(out["sum_price"][j]), int(out["sum_disc"][j]),
               int(out["sum_disc_price"][j]), out["sum_charge"][j], int(out["count"][j])]
        assert got == e


def test_q1_int128_charge_beyond_int64(orc):
    n = 64
    t = {"date": np.zeros(n, np.int32), "flag1": np.zeros(n, np.uint8), "flag2": np.zeros(n, np.uint8),
         "quantity": np.ones(n, np.int64), "price": np.full(n, 90_000_000_000_000_000, np.int64),
         "discount": np.zeros(n, np.int64), "tax": np.full(n, 8, np.int64)}
    out = orc.q1_flags(t, 0)
    assert out["sum_charge"][0] == n * 90_000_000_000_000_000 * 100 * 108
    assert out["sum_charge"][0] > 2**63


def test_q1_config_flags_six_groups(orc):
    t = orc.fact_table(100_000)
    out = orc.q1_flags(t, 2505)
    assert list(out["key"]) == [0, 1, 256, 257, 512, 513]
    assert int(out["count"].sum()) == int(np.count_nonzero(t["date"] <= 2505))
    assert int(out["sum_qty"].sum()) == int(t["quantity"][t["date"] <= 2505].sum())


def test_select_int64_keys(orc):
    k = orc.gen_range(np.int64, 5, 1, 30_001, -(1 << 40), 1 << 40)
    p = orc.gen_u32(5, 2, 30_001)
    op, orow = orc.select(k, p, -5, 1 << 39)
    idx = [i for i in range(k.size) if -5 <= int(k[i]) < (1 << 39)]
    assert list(orow) == idx and np.array_equal(op, p[idx])


# ---------------- TPC-H Q6 shape ----------------

def _q6_brute(t, lo, hi, dlo, dhi, qlt):
    rev, cnt = 0, 0
    for i in range(t["date"].size):
        if lo <= int(t["date"][i]) < hi and int(t["quantity"][i]) < qlt and dlo <= int(t["discount"][i]) <= dhi:
            rev += int(t["price"][i]) * int(t["discount"][i])
            cnt += 1
    rev %= 1 << 64
    return rev - (1 << 64) if rev >= 1 << 63 else rev, cnt


@pytest.mark.parametrize("n", [0, 1, 7999, 8001, 30_011])
def test_q6_matches_python_integers(orc, n):
    t = orc.fact_table(n, seed=71 + n)
    assert orc.q6(t) == _q6_brute(t, 731, 1096, 5, 7, 24)


def test_q6_bounds_are_half_open_date_inclusive_discount_strict_quantity(orc):
    # one row on each side of every bound: only the rows inside all three count
    date = np.array([730, 731, 1095, 1096, 900, 900, 900, 900, 900, 900], np.int32)
    disc = np.array([6, 6, 6, 6, 4, 5, 7, 8, 6, 6], np.int64)
    qty = np.array([10, 10, 10, 10, 10, 10, 10, 10, 23, 24], np.int64)
    price = np.arange(1, 11, dtype=np.int64) * 1000
    t = {"date": date, "quantity": qty, "price": price, "discount": disc}
    keep = [1, 2, 5, 6, 8]
    assert orc.q6(t) == (sum(int(price[i]) * int(disc[i]) for i in keep), len(keep))


def test_q6_revenue_wraps_like_int64(orc):
    n = 64
    t = {"date": np.full(n, 800, np.int32), "quantity": np.ones(n, np.int64),
         "price": np.full(n, (1 << 62) - 5, np.int64), "discount": np.full(n, 7, np.int64)}
    assert orc.q6(t) == _q6_brute(t, 731, 1096, 5, 7, 24)


def test_q6_matches_numpy_at_config_size_sample(orc):
    t = orc.fact_table(2_000_000)
    m = ((t["date"] >= 731) & (t["date"] < 1096) & (t["quantity"] < 24) & (t["discount"] >= 5) & (t["discount"] <= 7))
    rev = int((t["price"][m].astype(np.uint64) * t["discount"][m].astype(np.uint64)).sum(dtype=np.uint64))
    assert orc.q6(t) == (rev - (1 << 64) if rev >= 1 << 63 else rev, int(m.sum()))


# ---------------- packed dates ----------------
# pimdal/main/tpc_h/reader/read_table.cpp:8-16: the calendar fields come from
# the C library (strptime / gmtime_r) in the oracle; Python's datetime is an
# independent calendar.  The packing expression is the reference's own
# (day + (month << 5) + ((year - 1900) << 9)); its decode below recovers the
# fields, so a slip in a shift or in the month base fails.

def _py_pack(d):
    return d.day + (d.month << 5) + ((d.year - 1900) << 9)


def _unpack(p):
    p = int(p)
    return (p >> 9) + 1900, (p >> 5) & 15, p & 31


def test_date32_pack_matches_python_calendar(orc):
    import datetime as dt
    epoch = dt.date(1970, 1, 1)
    lo, hi = (dt.date(1, 1, 1) - epoch).days, (dt.date(9999, 12, 31) - epoch).days
    rng = np.random.default_rng(11)
    days = np.concatenate([rng.integers(lo, hi + 1, 20000), np.arange(-800, 800),
                           [lo, hi, 0, -1, 1, (dt.date(1900, 2, 28) - epoch).days, (dt.date(1900, 3, 1) - epoch).days,
                            (dt.date(2000, 2, 29) - epoch).days, (dt.date(2100, 3, 1) - epoch).days]]).astype(np.int32)
    got = orc.date32_pack(days)
    for x, p in zip(days.tolist(), got.tolist()):
        d = epoch + dt.timedelta(days=x)
        assert p == _py_pack(d), (x, d, p)
        assert _unpack(p) == (d.year, d.month, d.day)


def test_date32_pack_is_order_preserving(orc):
    # the TPC-H predicates compare packed dates: later date, larger value
    days = np.arange(-30000, 60000, dtype=np.int32)
    assert np.all(np.diff(orc.date32_pack(days).astype(np.int64)) > 0)


def test_date_pack_str_reference_literals(orc):
    # the literals of pimdal/main/tpc_h/query{1,4,5,6}/host/host_q*.cpp
    import datetime as dt
    for s in ["1998-09-02", "1993-07-01", "1993-10-01", "1993-08-01", "1993-11-01", "1994-01-01", "1995-01-01"]:
        assert orc.date_pack_str(s) == _py_pack(dt.date.fromisoformat(s))
    assert orc.date_pack_str("not a date") is None


# ---------------- Bloom-filter semi-join ----------------
# pimdal/join/hash_join.c:117-152, 237-341; pimdal/hash/hash_steps.c:143-183.
# The mixers and the filter bit positions are pinned against the reference's
# own hash_func.h in tests/test_ref_pins.py.  Here, what the mathematics
# fixes: no false negatives, a false-positive rate near (1 - e^(-4n/m))^4, a
# build key only touches its own partition's filter, at most 4 bits per key.

def test_bloom_no_false_negatives_and_fp_rate(orc):
    rng = np.random.default_rng(21)
    keys = np.unique(rng.integers(0, 1 << 32, 2000, dtype=np.uint64)).astype(np.uint32)
    filt = orc.bloom_build(keys, 0, 32)
    assert int(np.unpackbits(filt).sum()) <= 4 * keys.size
    probe = rng.integers(0, 1 << 32, 200_000, dtype=np.uint64).astype(np.uint32)
    probe = probe[~np.isin(probe, keys)]
    _, rows = orc.bloom_probe(np.concatenate([keys, probe]), None, 0, 32, filt)
    assert np.array_equal(rows[: keys.size], np.arange(keys.size))        # every build key is kept
    fp = (rows.size - keys.size) / probe.size
    theory = (1 - np.exp(-4 * keys.size / 16384)) ** 4
    assert 0.5 * theory < fp < 2 * theory, (fp, theory)


def test_bloom_partitions_are_isolated(orc):
    part_bits, shift = 4, 20
    rng = np.random.default_rng(22)
    keys = rng.integers(0, 1 << 32, 50_000, dtype=np.uint64).astype(np.uint32)
    part = np.array([(orc.pim_hash(0, int(k)) >> (shift - part_bits)) & 15 for k in keys[:4000]])
    sel = keys[:4000][part == 7]
    filt = orc.bloom_build(sel, part_bits, shift).reshape(16, 2048)
    assert filt[7].any() and not np.delete(filt, 7, axis=0).any()
    # probes that fall in another partition see an empty filter: nothing kept
    other = keys[:4000][part != 7]
    assert orc.bloom_probe(other, None, part_bits, shift, filt.ravel())[1].size == 0


# ---------------- TPC-H Q3 shape ----------------

def _q3_brute(t, segment, date):
    """The SQL of the Q3 shape by nested loops over the rows (tiny inputs):
    customers of the segment, their orders before `date`, those orders'
    lineitems shipped after `date`; revenue summed per (orderkey,
    orderdate, shippriority), ordered by revenue desc, orderdate, orderkey."""
    groups = {}
    for ci in range(len(t["c_custkey"])):
        if int(t["c_mktsegment"][ci]) != segment:
            continue
        for oi in range(len(t["o_orderkey"])):
            if int(t["o_custkey"][oi]) != int(t["c_custkey"][ci]) or not int(t["o_orderdate"][oi]) < date:
                continue
            for li in range(len(t["l_orderkey"])):
                if int(t["l_orderkey"][li]) != int(t["o_orderkey"][oi]) or not int(t["l_shipdate"][li]) > date:
                    continue
                g = (int(t["o_orderkey"][oi]), int(t["o_orderdate"][oi]), int(t["o_shippriority"][oi]))
                groups[g] = groups.get(g, 0) + int(t["l_extendedprice"][li]) * (100 - int(t["l_discount"][li]))
    rows = sorted(groups.items(), key=lambda kv: (-kv[1], kv[0][1], kv[0][0] & 0xFFFFFFFF))
    return rows


def _q3_tiny(seed):
    rng = np.random.default_rng(seed)
    nc, no, nl = 12, 30, 90
    t = {"c_custkey": rng.permutation(nc).astype(np.int32) * 3 - 5, "c_mktsegment": rng.integers(0, 3, nc).astype(np.uint8)}
    t["o_orderkey"] = (rng.permutation(no) * 7 - 40).astype(np.int32)
    t["o_custkey"] = t["c_custkey"][rng.integers(0, nc, no)]
    t["o_custkey"][:3] = 9999                                   # no such customer
    t["o_orderdate"] = rng.integers(100, 200, no).astype(np.int32)
    t["o_shippriority"] = rng.integers(0, 3, no).astype(np.int32)
    t["l_orderkey"] = t["o_orderkey"][rng.integers(0, no, nl)]
    t["l_orderkey"][:4] = 12345                                  # no such order
    t["l_extendedprice"] = rng.integers(-50, 1000, nl).astype(np.int64)
    t["l_discount"] = rng.integers(0, 11, nl).astype(np.int64)
    t["l_shipdate"] = rng.integers(120, 260, nl).astype(np.int32)
    return t


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("segment,date", [(0, 150), (1, 170), (2, 130), (7, 150)])
def test_q3_oracle_matches_brute_force(orc, seed, segment, date):
    t = _q3_tiny(seed)
    exp = _q3_brute(t, segment, date)
    got, g = orc.q3(t, segment, date)
    assert g == len(exp)
    assert [(int(k), int(d), int(p)) for k, d, p in zip(got["key"], got["orderdate"], got["shippriority"])] == \
        [e[0] for e in exp]
    assert [int(r) for r in got["revenue"]] == [e[1] for e in exp]


def test_q3_oracle_matches_pandas(orc):
    """Medium tables: the oracle against pandas' merges and group-by (a
    library's hash joins, as the reference checks its result with Acero,
    pimdal/main/tpc_h/query3/host/host_q3.cpp:350-420)."""
    import pandas as pd
    from pimdal_amd import datagen
    t = datagen.q3_tables(200_000, seed=5)
    date = int(datagen.pack_days(np.array([9204]))[0])          # 1995-03-15
    c = pd.DataFrame({"c_custkey": t["c_custkey"], "c_mktsegment": t["c_mktsegment"]})
    o = pd.DataFrame({k: t[k] for k in ("o_orderkey", "o_custkey", "o_orderdate", "o_shippriority")})
    li = pd.DataFrame({k: t[k] for k in ("l_orderkey", "l_extendedprice", "l_discount", "l_shipdate")})
    m = c[c.c_mktsegment == 1].merge(o[o.o_orderdate < date], left_on="c_custkey", right_on="o_custkey")
    m = m.merge(li[li.l_shipdate > date], left_on="o_orderkey", right_on="l_orderkey")
    m["rev"] = m.l_extendedprice * (100 - m.l_discount)
    g = m.groupby(["o_orderkey", "o_orderdate", "o_shippriority"], as_index=False)["rev"].sum()
    g = g.sort_values(["rev", "o_orderdate", "o_orderkey"], ascending=[False, True, True])
    got, n = orc.q3(t, 1, date)
    assert n == len(g) and n > 100
    assert np.array_equal(got["key"], g.o_orderkey.to_numpy()) and np.array_equal(got["revenue"], g.rev.to_numpy())
    assert np.array_equal(got["orderdate"], g.o_orderdate.to_numpy())
    assert np.array_equal(got["shippriority"], g.o_shippriority.to_numpy())


def test_q3_oracle_rejects_repeated_orderkeys(orc):
    t = _q3_tiny(9)
    t["o_orderkey"][5] = t["o_orderkey"][6]
    t["o_orderdate"][5] = t["o_orderdate"][6] = 100
    cust = t["c_custkey"][t["c_mktsegment"] == 0][0]
    t["o_custkey"][5] = t["o_custkey"][6] = cust
    with pytest.raises(ValueError):
        orc.q3(t, 0, 150)


def test_packed_dates_of_the_generator_match_the_oracle(orc):
    from pimdal_amd import datagen
    days = np.concatenate([np.arange(8035, 8035 + 2600), np.array([-719162, 0, 11016, 2932896])]).astype(np.int32)
    assert np.array_equal(datagen.pack_days(days), orc.date32_pack(days))


def _q4_brute(t, lo, hi):
    """The SQL of the Q4 shape by nested loops (tiny inputs): orders with
    lo <= o_orderdate < hi for which EXISTS a lineitem of that order with
    l_commitdate < l_receiptdate, counted per o_orderpriority, in priority
    order."""
    counts = {}
    for oi in range(len(t["o_orderkey"])):
        if not (lo <= int(t["o_orderdate"][oi]) < hi):
            continue
        if any(int(t["l_orderkey"][li]) == int(t["o_orderkey"][oi]) and
               int(t["l_commitdate"][li]) < int(t["l_receiptdate"][li]) for li in range(len(t["l_orderkey"]))):
            p = int(t["o_orderpriority"][oi])
            counts[p] = counts.get(p, 0) + 1
    return sorted(counts.items())


def _q4_tiny(seed):
    rng = np.random.default_rng(seed)
    no, nl = 40, 120
    t = {"o_orderkey": (rng.permutation(no) * 7 - 40).astype(np.int32),
         "o_orderdate": rng.integers(100, 200, no).astype(np.int32),
         "o_orderpriority": rng.integers(0, 5, no).astype(np.uint8)}
    t["l_orderkey"] = t["o_orderkey"][rng.integers(0, no, nl)]
    t["l_orderkey"][:4] = 12345                                  # no such order
    t["l_commitdate"] = rng.integers(100, 200, nl).astype(np.int32)
    t["l_receiptdate"] = rng.integers(100, 200, nl).astype(np.int32)
    t["l_receiptdate"][4:8] = t["l_commitdate"][4:8]             # equal dates do not qualify
    return t


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("lo,hi", [(100, 200), (120, 150), (150, 151), (300, 400)])
def test_q4_oracle_matches_brute_force(orc, seed, lo, hi):
    t = _q4_tiny(seed)
    exp = _q4_brute(t, lo, hi)
    codes, counts = orc.q4(t, lo, hi)
    assert [(int(c), int(n)) for c, n in zip(codes, counts)] == exp


def test_q4_oracle_matches_pandas(orc):
    """Medium tables: the oracle against pandas' semi-join and group-by."""
    import pandas as pd
    from pimdal_amd import datagen
    t = datagen.q4_tables(200_000, seed=6)
    lo, hi = (int(x) for x in datagen.pack_days(np.array([8582, 8674])))   # 1993-07-01, 1993-10-01
    o = pd.DataFrame({k: t[k] for k in ("o_orderkey", "o_orderdate", "o_orderpriority")})
    li = pd.DataFrame({k: t[k] for k in ("l_orderkey", "l_commitdate", "l_receiptdate")})
    keys = li[li.l_commitdate < li.l_receiptdate].l_orderkey.unique()
    q = o[(o.o_orderdate >= lo) & (o.o_orderdate < hi) & o.o_orderkey.isin(keys)]
    g = q.groupby("o_orderpriority").size().sort_index()
    codes, counts = orc.q4(t, lo, hi)
    assert len(g) == 5 and counts.sum() > 1000
    assert np.array_equal(codes, g.index.to_numpy().astype(np.uint8)) and np.array_equal(counts, g.to_numpy())


def test_q4_oracle_refuses_repeated_orderkeys(orc):
    t = _q4_tiny(1)
    t["o_orderkey"][1] = t["o_orderkey"][0]
    t["o_orderdate"][:2] = 150
    with pytest.raises(ValueError):
        orc.q4(t, 100, 200)


def _q5_brute(t, name, lo, hi):
    """The SQL of the Q5 shape by nested loops (tiny inputs): customer,
    orders, lineitem, supplier, nation, region joined on their keys, the
    supplier's nation equal to the customer's, the region named `name`,
    lo <= o_orderdate < hi; revenue per nation, descending (ties by key)."""
    region_keys = {int(k) for k, n in zip(t["r_regionkey"], t["r_name"]) if int(n) == name}
    nations = {int(k) for k, r in zip(t["n_nationkey"], t["n_regionkey"]) if int(r) in region_keys}
    rev = {}
    for ci in range(len(t["c_custkey"])):
        cn = int(t["c_nationkey"][ci])
        if cn not in nations:
            continue
        for oi in range(len(t["o_orderkey"])):
            if int(t["o_custkey"][oi]) != int(t["c_custkey"][ci]) or not lo <= int(t["o_orderdate"][oi]) < hi:
                continue
            for li in range(len(t["l_orderkey"])):
                if int(t["l_orderkey"][li]) != int(t["o_orderkey"][oi]):
                    continue
                for si in range(len(t["s_suppkey"])):
                    if int(t["s_suppkey"][si]) == int(t["l_suppkey"][li]) and int(t["s_nationkey"][si]) == cn:
                        rev[cn] = rev.get(cn, 0) + int(t["l_extendedprice"][li]) * (100 - int(t["l_discount"][li]))
    return sorted(rev.items(), key=lambda kv: (-kv[1], kv[0]))


def _q5_tiny(seed):
    rng = np.random.default_rng(seed)
    nc, no, nl, ns = 15, 40, 120, 8
    t = {"r_regionkey": np.array([3, 1, 2], np.int32), "r_name": np.array([0, 1, 2], np.uint8),
         "n_nationkey": np.arange(6, dtype=np.int32) * 2, "n_regionkey": np.array([3, 1, 2, 1, 3, 2], np.int32)}
    t["c_custkey"] = (rng.permutation(nc) * 3 - 5).astype(np.int32)
    t["c_nationkey"] = (rng.integers(0, 7, nc) * 2).astype(np.int32)          # 12: no such nation
    t["o_orderkey"] = (rng.permutation(no) * 7 - 40).astype(np.int32)
    t["o_custkey"] = t["c_custkey"][rng.integers(0, nc, no)]
    t["o_custkey"][:3] = 9999                                                 # no such customer
    t["o_orderdate"] = rng.integers(100, 200, no).astype(np.int32)
    t["l_orderkey"] = t["o_orderkey"][rng.integers(0, no, nl)]
    t["s_suppkey"] = (rng.permutation(ns) + 1).astype(np.int32)
    t["s_nationkey"] = (rng.integers(0, 6, ns) * 2).astype(np.int32)
    t["l_suppkey"] = t["s_suppkey"][rng.integers(0, ns, nl)]
    t["l_suppkey"][:4] = 555                                                  # no such supplier
    t["l_extendedprice"] = rng.integers(-50, 1000, nl).astype(np.int64)
    t["l_discount"] = rng.integers(0, 11, nl).astype(np.int64)
    return t


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("name,lo,hi", [(0, 100, 200), (1, 120, 170), (2, 100, 200), (2, 150, 150), (7, 100, 200)])
def test_q5_oracle_matches_brute_force(orc, seed, name, lo, hi):
    t = _q5_tiny(seed)
    exp = _q5_brute(t, name, lo, hi)
    nat, rev = orc.q5(t, name, lo, hi)
    assert [(int(a), int(b)) for a, b in zip(nat, rev)] == exp


def test_q5_oracle_matches_pandas(orc):
    """Medium tables: the oracle against pandas' merges (customer -> orders ->
    lineitem -> supplier on (suppkey, nation)) and group-by."""
    import pandas as pd
    from pimdal_amd import datagen
    t = datagen.q5_tables(400_000, seed=7)
    asia = datagen.Q5_REGIONS.index("ASIA")
    lo, hi = (int(x) for x in datagen.pack_days(np.array([8766, 9131])))   # 1994-01-01, 1995-01-01
    df = lambda *cols: pd.DataFrame({k: t[k] for k in cols})
    reg = df("r_regionkey", "r_name")
    nat = df("n_nationkey", "n_regionkey").merge(reg[reg.r_name == asia], left_on="n_regionkey", right_on="r_regionkey")
    cu = df("c_custkey", "c_nationkey").merge(nat, left_on="c_nationkey", right_on="n_nationkey")
    o = df("o_orderkey", "o_custkey", "o_orderdate")
    m = cu.merge(o[(o.o_orderdate >= lo) & (o.o_orderdate < hi)], left_on="c_custkey", right_on="o_custkey")
    m = m.merge(df("l_orderkey", "l_suppkey", "l_extendedprice", "l_discount"), left_on="o_orderkey",
                right_on="l_orderkey")
    m = m.merge(df("s_suppkey", "s_nationkey"), left_on=["l_suppkey", "c_nationkey"],
                right_on=["s_suppkey", "s_nationkey"])
    m["rev"] = m.l_extendedprice * (100 - m.l_discount)
    g = m.groupby("c_nationkey", as_index=False)["rev"].sum().sort_values(["rev", "c_nationkey"],
                                                                          ascending=[False, True])
    nat_o, rev_o = orc.q5(t, asia, lo, hi)
    assert len(g) == 5
    assert np.array_equal(nat_o, g.c_nationkey.to_numpy()) and np.array_equal(rev_o, g.rev.to_numpy())
