"""Selection on the GPU vs the CPU restatement: bit-exact payloads and row ids."""
import numpy as np
import pytest

from pimdal_amd import operators as ops, datagen

pytestmark = pytest.mark.gpu
T = datagen.SELECT_THRESHOLD


def run(ctx, k, p, lo, hi):
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    op, orow, cnt = ops.select(ctx, dk, dp, lo, hi)
    return op.numpy(cnt), orow.numpy(cnt)


@pytest.fixture(params=[(0, 256), (0, 1024), (1, 256)], ids=["one_pass", "one_pass_16k_tiles", "two_pass"],
                autouse=True)
def select_mode(request, ctx):
    """Every select test runs in each mode: one pass with decoupled look-back
    over 4096-row tiles and over 16384-row tiles, and count pass + write
    pass."""
    mode, block = request.param
    ctx.set_tune("select_mode", mode)
    ctx.set_tune("select_block", block)
    yield request.param
    ctx.set_tune("select_mode", 0)
    ctx.set_tune("select_block", 1024)


def test_select_2_20_uniform_10pct(ctx, orc):
    k, p = orc.select_workload(1 << 20)
    gp, gr = run(ctx, k, p, 0, T)
    ep, er = orc.select(k, p, 0, T)
    assert np.array_equal(gp, ep) and np.array_equal(gr, er)


@pytest.mark.parametrize("n", [0, 1, 63, 64, 255, 4095, 4096, 4097, 100_003])
@pytest.mark.parametrize("lo,hi", [(0, T), (0, 0), (0, 1 << 32), (1 << 31, 3 << 30)])
def test_select_edge_sizes(ctx, orc, n, lo, hi):
    k, p = orc.select_workload(n, seed=500 + n)
    gp, gr = run(ctx, k, p, lo, hi)
    ep, er = orc.select(k, p, lo, min(hi, (1 << 32) - 1)) if hi <= (1 << 32) - 1 else (p, np.arange(n, dtype=np.uint32))
    assert np.array_equal(gp, ep) and np.array_equal(gr, er)


def test_select_int32_reference_predicate(ctx, orc):
    k = orc.gen_range(np.int32, 8, 1, 300_001, -40, 50)
    p = orc.gen_u32(8, 2, 300_001)
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    op, orow, cnt = ops.select(ctx, dk, dp, 10, 30)
    ep, er = orc.select(k, p, 10, 30)
    assert np.array_equal(op.numpy(cnt), ep) and np.array_equal(orow.numpy(cnt), er)


def test_select_unaligned_input(ctx, orc):
    k, p = orc.select_workload(50_001, seed=3)
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    vk, vp = dk.view(1, 50_000), dp.view(3, 49_998)
    vk = dk.view(1, 49_998)
    op, orow, cnt = ops.select(ctx, vk, vp, 0, T)
    ep, er = orc.select(k[1:49_999], p[3:50_001], 0, T)
    assert np.array_equal(op.numpy(cnt), ep) and np.array_equal(orow.numpy(cnt), er)


def test_select_2_27_full(ctx, orc):
    n = 1 << 27
    dk, dp = datagen.select_workload(ctx, n)
    op, orow, cnt = ops.select(ctx, dk, dp, 0, T)
    k, p = orc.select_workload(n)
    ep, er = orc.select(k, p, 0, T)
    assert cnt == len(ep)
    assert np.array_equal(op.numpy(cnt), ep) and np.array_equal(orow.numpy(cnt), er)


@pytest.mark.parametrize("n", [1, 4097, 300_003])
def test_select_int64_keys(ctx, orc, n):
    k = orc.gen_range(np.int64, 6, 1, n, 90000, 10500000)   # the fact table's price column shape
    p = orc.gen_u32(6, 2, n)
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    for lo, hi in ((0, 1_000_000), (-(1 << 62), 5_000_000), (10500000, 1 << 40)):
        op, orow, cnt = ops.select(ctx, dk, dp, lo, hi)
        ep, er = orc.select(k, p, lo, hi)
        assert np.array_equal(op.numpy(cnt), ep) and np.array_equal(orow.numpy(cnt), er)


def brute_select(k, p, lo, hi):
    """lo <= key < hi in exact integer arithmetic (bounds may lie outside the
    key type, which the C ABI's int64 bounds allow)."""
    ko = k.astype(object)
    rows = np.nonzero((ko >= lo) & (ko < hi))[0].astype(np.uint32)
    return p[rows], rows


@pytest.mark.parametrize("lo,hi", [(-(1 << 40), 1 << 40), (-(1 << 31), -(1 << 31) + 1), ((1 << 31) - 1, 1 << 31),
                                   (-5, 5), (7, 7), (10, 3), (-(1 << 63), -(1 << 31)), (0, 1 << 62)])
def test_select_int32_full_domain_bounds(ctx, orc, lo, hi):
    k = orc.gen_u32(9, 1, 200_003).view(np.int32).copy()
    k[:4] = [-(1 << 31), (1 << 31) - 1, 0, -1]
    p = orc.gen_u32(9, 2, 200_003)
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    op, orow, cnt = ops.select(ctx, dk, dp, lo, hi)
    ep, er = brute_select(k, p, lo, hi)
    assert np.array_equal(op.numpy(cnt), ep) and np.array_equal(orow.numpy(cnt), er)


@pytest.mark.parametrize("lo,hi", [(-(1 << 63), (1 << 63) - 1), (-(1 << 63), 0), (0, (1 << 63) - 1), (-1, 1),
                                   (1 << 62, -(1 << 62))])
def test_select_int64_full_domain_bounds(ctx, orc, lo, hi):
    n = 100_001
    k = ((orc.gen_u32(10, 1, n).astype(np.uint64) << np.uint64(32)) | orc.gen_u32(10, 2, n)).view(np.int64).copy()
    k[:4] = [-(1 << 63), (1 << 63) - 1, 0, -1]
    p = orc.gen_u32(10, 3, n)
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    op, orow, cnt = ops.select(ctx, dk, dp, lo, hi)
    ep, er = brute_select(k, p, lo, hi)
    assert np.array_equal(op.numpy(cnt), ep) and np.array_equal(orow.numpy(cnt), er)
    if -(1 << 63) <= lo and hi <= (1 << 63) - 1:   # inside int64: the oracle agrees too
        op2, or2 = orc.select(k, p, lo, hi)
        assert np.array_equal(or2, er)


@pytest.mark.parametrize("n", [0, 1, 3, 1021, 1024, 4097, 100_003])
@pytest.mark.parametrize("lo,hi", [(1, 2), (0, 256), (-5, 7), (200, 1000), (300, 400)])
@pytest.mark.parametrize("with_payload", [True, False])
def test_select_uint8_keys(ctx, orc, n, lo, hi, with_payload):
    """uint8 predicate columns (the flag columns of the fact table, the Bloom
    probe's keep flags): the same rows as the oracle's uint32 selection over
    the same values with the range clamped to [0, 256); ragged row counts
    (within a dword, within a wave) and no payload column (row ids only)."""
    rng = np.random.default_rng(700 + n)
    k = rng.integers(0, 256, n, dtype=np.uint16).astype(np.uint8)
    p = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    ep, er = orc.select(k.astype(np.uint32), p, max(lo, 0), min(hi, 256))
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    op, orow, cnt = ops.select(ctx, dk, dp if with_payload else None, lo, hi, want_payload=with_payload)
    assert cnt == er.size
    assert np.array_equal(orow.numpy(cnt), er)
    if with_payload:
        assert np.array_equal(op.numpy(cnt), ep)


def test_select_uint8_unaligned(ctx, orc):
    """A uint8 column starting one byte into its allocation (dword loads off)."""
    rng = np.random.default_rng(5)
    k = rng.integers(0, 256, 70_001, dtype=np.uint16).astype(np.uint8)
    p = rng.integers(0, 1 << 32, 70_001, dtype=np.uint64).astype(np.uint32)
    dk, dp = ctx.to_device(k), ctx.to_device(p)
    vk, vp = dk.view(1, 70_000), dp.view(1, 70_000)
    op, orow, cnt = ops.select(ctx, vk, vp, 10, 90)
    ep, er = orc.select(k[1:].astype(np.uint32), p[1:], 10, 90)
    assert np.array_equal(op.numpy(cnt), ep) and np.array_equal(orow.numpy(cnt), er)
