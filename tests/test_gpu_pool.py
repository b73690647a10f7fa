"""The context's caching device allocator: budget accounting, best-fit reuse
of cached blocks, and trimming of the cache under a fixed budget."""
import numpy as np
import pytest

import pimdal_amd

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def test_cached_block_reused_within_an_eighth():
    """Sizes round up to 2 MiB steps above 1 MiB; a cached block up to 1/8
    larger than the rounded request is taken instead of a new one."""
    with pimdal_amd.Context(0) as c:
        a = c.empty(20 * MiB, np.uint8)
        a.free()
        c.sync()
        before = c.mem_stats()
        b = c.empty(18 * MiB, np.uint8)                    # a smaller rounded size, within 1/8 of the cached block
        after = c.mem_stats()
        assert after["reserved"] == before["reserved"]   # the cached 20 MiB block was reused, nothing new reserved
        b.free()
        big = c.empty(24 * MiB, np.uint8)                  # larger than every cached block: a new one
        assert c.mem_stats()["reserved"] > after["reserved"]
        big.free()


def test_budget_holds_over_many_sizes():
    budget = 256 * MiB
    with pimdal_amd.Context(0, budget) as c:
        rng = np.random.default_rng(4)
        for _ in range(400):
            n = int(rng.integers(1, 64)) * MiB + int(rng.integers(0, 4096))
            a = c.empty(n, np.uint8)
            st = c.mem_stats()
            assert st["reserved"] <= budget and st["live"] <= st["reserved"]
            a.free()
        with pytest.raises(pimdal_amd.PimdalError):
            c.empty(budget + MiB, np.uint8)                # larger than the whole budget
        st = c.mem_stats()
        assert st["live"] == 0 and st["budget"] == budget


def test_lowering_the_budget_gives_cached_memory_back():
    """Reserved bytes (live + cached) stay within the budget after it is
    lowered: cached blocks are freed at once, live ones stay."""
    with pimdal_amd.Context(0, 512 * MiB) as c:
        keep = c.empty(64 * MiB, np.uint8)
        for n in (100, 80, 60):
            a = c.empty(n * MiB, np.uint8)
            a.free()
        c.sync()
        st = c.mem_stats()
        assert st["reserved"] >= (64 + 100 + 80 + 60) * MiB
        c.set_budget(128 * MiB)
        st = c.mem_stats()
        assert st["budget"] == 128 * MiB and st["reserved"] <= 128 * MiB and st["live"] == 64 * MiB
        keep.free()
