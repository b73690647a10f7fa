"""The Bloom-filter semi-join (pimdal/join/hash_join.c:117-152, 237-341) on
the GPU against the CPU restatement: filters byte for byte, kept rows in
row order, on the LDS path (up to 32 filters) and the HBM path."""
import numpy as np
import pytest

import pimdal_amd
from pimdal_amd import operators as ops

pytestmark = pytest.mark.gpu


def _keys(seed, n):
    return np.random.default_rng(seed).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("part_bits,shift", [(0, 32), (3, 32), (5, 17), (6, 32), (10, 24)])
@pytest.mark.parametrize("nb", [0, 1, 1000, 1 << 20])
def test_bloom_filters_match_oracle(ctx, orc, part_bits, shift, nb):
    keys = _keys(nb + part_bits, nb)
    filt = ops.bloom_build(ctx, ctx.to_device(keys), part_bits, shift).numpy()
    assert np.array_equal(filt.view(np.uint8), orc.bloom_build(keys, part_bits, shift))


@pytest.mark.parametrize("part_bits,shift", [(0, 32), (5, 32), (6, 32), (7, 20), (9, 30), (10, 24)])
@pytest.mark.parametrize("np_", [0, 1, 4097, 1 << 21])
@pytest.mark.parametrize("with_payload", [True, False])
def test_bloom_probe_matches_oracle(ctx, orc, part_bits, shift, np_, with_payload):
    build = _keys(7, 3000 << part_bits)
    probe = _keys(8, np_)
    probe[::3] = build[np.arange(probe[::3].size) % build.size]            # a third are members
    pay = _keys(9, np_) if with_payload else None
    f = orc.bloom_build(build, part_bits, shift)
    ep, er = orc.bloom_probe(probe, pay, part_bits, shift, f)
    df = ctx.to_device(f.view(np.uint32))
    gp, gr, cnt = ops.bloom_probe(ctx, ctx.to_device(probe), ctx.to_device(pay) if with_payload else None,
                                  part_bits, shift, df)
    assert cnt == ep.size
    assert np.array_equal(gp.numpy(cnt), ep) and np.array_equal(gr.numpy(cnt), er)
    sp, sr, scnt = ops.bloom_semi_join(ctx, ctx.to_device(build), ctx.to_device(probe),
                                       ctx.to_device(pay) if with_payload else None, part_bits, shift)
    assert scnt == cnt and np.array_equal(sp.numpy(scnt), ep) and np.array_equal(sr.numpy(scnt), er)


def test_bloom_rejects_bad_partitions(ctx):
    k = ctx.to_device(_keys(1, 10))
    for bits, shift in [(17, 32), (5, 4), (3, 33)]:
        with pytest.raises(pimdal_amd.PimdalError) as e:
            ops.bloom_build(ctx, k, bits, shift) if bits <= 16 else ops.bloom_semi_join(ctx, k, k, None, bits, shift)
        assert e.value.code == -1


@pytest.mark.parametrize("part_bits,shift", [(6, 32), (8, 27), (9, 32), (10, 24)])
@pytest.mark.parametrize("np_", [1 << 16, (1 << 16) + 1, (1 << 22) + 12345])
def test_bloom_filter_groups_match_oracle(ctx, orc, part_bits, shift, np_):
    """The probe through filter groups held in LDS (64 filters a group, a
    block per group streaming every chunk; 16384- and 32768-row chunks, one
    or two chunks loading ahead) and the probe reading the filters through
    L2 give the oracle's kept rows, in row order; sizes cover a ragged last
    chunk and one row past the group path's minimum."""
    build = _keys(17 + part_bits, 2000 << part_bits)
    probe = _keys(18, np_)
    probe[::4] = build[np.arange(probe[::4].size) % build.size]
    pay = _keys(19, np_)
    f = orc.bloom_build(build, part_bits, shift)
    ep, er = orc.bloom_probe(probe, pay, part_bits, shift, f)
    df, dk, dp = ctx.to_device(f.view(np.uint32)), ctx.to_device(probe), ctx.to_device(pay)
    try:
        for groups, shape in ((1, 0), (1, 1), (1, 2), (0, 0)):
            ctx.set_tune("bloom_groups", groups)
            ctx.set_tune("bloom_group_shape", shape)
            gp, gr, cnt = ops.bloom_probe(ctx, dk, dp, part_bits, shift, df)
            assert cnt == ep.size, (groups, shape)
            assert np.array_equal(gp.numpy(cnt), ep) and np.array_equal(gr.numpy(cnt), er), (groups, shape)
    finally:
        ctx.clear_tune("bloom_groups")
        ctx.clear_tune("bloom_group_shape")
