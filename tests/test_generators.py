"""Synthetic-data generators: bijectivity, ranges, distribution and fixed
digests (tests/golden/gen_digests.json, written by tests/golden/make_digests.py)."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "gen_digests.json"


@pytest.mark.parametrize("m", [1, 2, 3, 1000, 1 << 16, 100_003])
def test_perm_is_a_bijection(orc, m):
    p = orc.gen_perm(42, 3, m, m)
    assert p.max() < m and np.unique(p).size == m


def test_perm_offset_consistent(orc):
    a = orc.gen_perm(5, 3, 5000, 5000)
    b = orc.gen_perm(5, 3, 5000, 1000, offset=2000)
    assert np.array_equal(a[2000:3000], b)


def test_ranges_in_bounds(orc):
    t = orc.fact_table(100_000)
    assert t["quantity"].min() >= 1 and t["quantity"].max() <= 50
    assert t["price"].min() >= 90000 and t["price"].max() <= 10500000
    assert t["discount"].max() <= 10 and t["tax"].max() <= 8 and t["date"].max() <= 2556
    assert set(np.unique(t["flag1"])) == {0, 1, 2} and set(np.unique(t["flag2"])) == {0, 1}


def test_orderkey_groups_repeat_four_times(orc):
    ok = orc.gen_orderkey(9, 40_000, 10_000)
    u, c = np.unique(ok, return_counts=True)
    assert u.size == 10_000 and np.all(c == 4) and ok.min() >= 0


def test_zipf_frequencies(orc):
    m, n = 1000, 400_000
    r = orc.gen_zipf(3, 7, m, 1.0, n)
    assert r.min() >= 1 and r.max() <= m
    H = sum(1.0 / k for k in range(1, m + 1))
    f1 = np.count_nonzero(r == 1) / n
    f2 = np.count_nonzero(r == 2) / n
    assert abs(f1 - 1 / H) < 0.005 and abs(f2 - 0.5 / H) < 0.004


def test_zipf_top_key_share_2_24_distinct(orc):
    # hottest of 2^24 values holds about 1/H(2^24) = 5.8 % of rows
    r = orc.gen_zipf(1, 7, 1 << 24, 1.0, 200_000)
    assert abs(np.count_nonzero(r == 1) / 200_000 - 0.0581) < 0.004


def test_fmix32_bijective_on_sample(orc):
    a = np.arange(1 << 20, dtype=np.uint32)
    assert np.unique(orc.fmix32(a)).size == a.size


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_generator_digests_fixed(orc):
    want = json.loads(GOLDEN.read_text())
    from tests.golden.make_digests import compute
    assert compute(orc) == want
