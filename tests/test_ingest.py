"""Arrow / Parquet ingest and egress, and packed dates
(pimdal/main/tpc_h/reader/read_table.cpp): the host-side date functions of
the C ABI against the oracle on the CPU; chunked Arrow import, date32
conversion on the device and the Parquet reader on the GPU."""
import numpy as np
import pytest

import pimdal_amd
from pimdal_amd import operators as ops, ingest


# ---------------- host-side packing (no device work) ----------------

@pytest.mark.parametrize("text", ["1998-09-02", "1994-01-01", "1900-01-01", "2000-02-29", "1899-12-31", "0001-01-01",
                                  "9999-12-31", " 1995-1-5 ", "1998-02-30"])
def test_date_pack_str_matches_oracle(orc, text):
    assert ops.date_pack_str(text) == orc.date_pack_str(text)


@pytest.mark.parametrize("text", ["", "1998", "1998-09", "1998/09/02", "98-9-2x", "1998-13-01", "1998-00-10",
                                  "1998-01-32", "1998-01-00", "19980-01-01"])
def test_date_pack_str_rejects_malformed(text):
    with pytest.raises(pimdal_amd.PimdalError) as e:
        ops.date_pack_str(text)
    assert e.value.code == -1


def test_date_pack_numbers_match_strings(orc):
    rng = np.random.default_rng(5)
    for y, m, d in zip(rng.integers(1, 9999, 300), rng.integers(1, 13, 300), rng.integers(1, 29, 300)):
        assert ops.date_pack(int(y), int(m), int(d)) == orc.date_pack_str(f"{y:04d}-{m:02d}-{d:02d}")
    with pytest.raises(pimdal_amd.PimdalError):
        ops.date_pack(1998, 0, 1)


# ---------------- device ----------------

@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 3, 4, 1023, 1 << 20])
def test_date32_pack_matches_oracle(ctx, orc, n):
    rng = np.random.default_rng(n)
    days = rng.integers(-(1 << 31), (1 << 31) - 1, n, dtype=np.int64).astype(np.int32)
    days[: min(n, 8)] = np.array([-(1 << 31), (1 << 31) - 1, 0, -1, -25567, 10471, 11016, 47482], np.int32)[: min(n, 8)]
    d = ctx.to_device(days)
    got = ops.date32_pack(ctx, d).numpy()
    assert np.array_equal(got, orc.date32_pack(days))


@pytest.mark.gpu
def test_date32_pack_in_place_and_unaligned(ctx, orc):
    days = np.arange(-40000, 40000, 7, dtype=np.int32)
    exp = orc.date32_pack(days)
    d = ctx.to_device(days)
    ops.date32_pack(ctx, d, d)                      # in place, 16-byte aligned
    assert np.array_equal(d.numpy(), exp)
    d2 = ctx.to_device(days)
    v = d2.view(1, days.size - 1)                   # 4-byte aligned: the scalar path
    out = ctx.empty(days.size - 1, np.int32)
    ops.date32_pack(ctx, v, out.view(0, days.size - 1))
    assert np.array_equal(out.numpy(), exp[1:])


@pytest.mark.gpu
def test_import_chunked_array(ctx):
    import pyarrow as pa
    rng = np.random.default_rng(3)
    parts = [rng.integers(0, 1 << 32, m, dtype=np.uint64).astype(np.uint32) for m in (1000, 0, 1, 77777)]
    sliced = pa.array(parts[3]).slice(5, 70000)     # a chunk with an offset
    ca = pa.chunked_array([pa.array(parts[0]), pa.array(parts[1]), pa.array(parts[2]), sliced])
    col = ingest.import_arrow(ctx, ca)
    assert col.dtype == np.uint32 and col.n == 1000 + 0 + 1 + 70000
    assert np.array_equal(col.numpy(), np.concatenate([parts[0], parts[1], parts[2], parts[3][5:70005]]))
    back = ingest.to_arrow(ctx, col)
    assert back.type == pa.uint32() and np.array_equal(back.to_numpy(), col.numpy())
    col.free()


@pytest.mark.gpu
def test_import_refuses_nulls_and_other_types(ctx):
    import pyarrow as pa
    ok = pa.array(np.arange(10, dtype=np.int64))
    with pytest.raises(pimdal_amd.PimdalError) as e:
        ingest.import_arrow(ctx, pa.chunked_array([ok, pa.array([1, None, 3], type=pa.int64())]))
    assert e.value.code == -2
    with pytest.raises(pimdal_amd.PimdalError) as e:
        ingest.import_arrow(ctx, pa.array(["a", "b"]))
    assert e.value.code == -2


@pytest.mark.gpu
def test_read_parquet_row_groups_and_dates(ctx, orc, tmp_path):
    import pyarrow as pa
    import pyarrow.parquet as pq
    n = 300_001
    t = orc.fact_table(n)
    epoch_1992 = 8035                                # 1992-01-01 in days since 1970-01-01
    ship = (epoch_1992 + t["date"]).astype(np.int32)
    table = pa.table({"l_orderkey": pa.array(t["orderkey"].view(np.uint32)), "l_quantity": t["quantity"],
                      "l_shipdate": pa.array(ship, type=pa.date32()), "l_returnflag": t["flag1"]})
    path = tmp_path / "lineitem.parquet"
    pq.write_table(table, path, row_group_size=65536)
    assert pq.ParquetFile(path).num_row_groups == 5
    cols = ingest.read_parquet(ctx, str(path), columns=[0, 2, 3])          # by index, as read_table.cpp
    assert sorted(cols) == ["l_orderkey", "l_returnflag", "l_shipdate"]
    assert np.array_equal(cols["l_orderkey"].numpy(), t["orderkey"].view(np.uint32))
    assert np.array_equal(cols["l_returnflag"].numpy(), t["flag1"])
    assert np.array_equal(cols["l_shipdate"].numpy(), orc.date32_pack(ship))
    raw = ingest.read_parquet(ctx, str(path), columns=["l_shipdate", "l_quantity"], pack_dates=False)
    assert np.array_equal(raw["l_shipdate"].numpy(), ship)
    assert np.array_equal(raw["l_quantity"].numpy(), t["quantity"])
    # a packed literal compares with the packed column: the Q6 date range
    lo, hi = ops.date_pack_str("1994-01-01"), ops.date_pack_str("1995-01-01")
    packed = cols["l_shipdate"].numpy()
    import datetime as dt
    ref = np.array([dt.date(1994, 1, 1) <= dt.date(1970, 1, 1) + dt.timedelta(days=int(x)) < dt.date(1995, 1, 1)
                    for x in ship[:20000]])
    assert np.array_equal((packed[:20000] >= lo) & (packed[:20000] < hi), ref)
    for c in list(cols.values()) + list(raw.values()):
        c.free()
