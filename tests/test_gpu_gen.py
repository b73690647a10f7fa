"""Device generators produce exactly the host restatement's columns."""
import numpy as np
import pytest

from pimdal_amd import datagen

pytestmark = pytest.mark.gpu


def test_select_and_sort_workloads(ctx, orc):
    n = 100_003
    a, b = datagen.select_workload(ctx, n)
    ea, eb = orc.select_workload(n)
    assert np.array_equal(a.numpy(), ea) and np.array_equal(b.numpy(), eb)
    for z in (False, True):
        k, v = datagen.sort_workload(ctx, n, zipf=z)
        ek, ev = orc.sort_workload(n, zipf=z)
        assert np.array_equal(k.numpy(), ek) and np.array_equal(v.numpy(), ev)


@pytest.mark.parametrize("zipf", [False, True])
def test_join_workload(ctx, orc, zipf):
    d = datagen.join_workload(ctx, 1 << 16, 1 << 18, zipf=zipf)
    e = orc.join_workload(1 << 16, 1 << 18, zipf=zipf)
    assert all(np.array_equal(x.numpy(), y) for x, y in zip(d, e))


def test_fact_table(ctx, orc):
    d = datagen.fact_table(ctx, 60_000)
    e = orc.fact_table(60_000)
    for c in e:
        assert np.array_equal(d[c].numpy(), e[c]), c
