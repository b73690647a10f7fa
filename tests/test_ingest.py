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
def test_import_keeps_nulls_and_refuses_other_types(ctx):
    """Nullable chunks (with offsets) import with their validity concatenated
    at bit level and export back unchanged; strings are refused."""
    import pyarrow as pa
    vals = [1, None, 3, None, 5, 6, 7, None, 9, 10, None]
    a = pa.array(vals, type=pa.int64())
    ch = pa.chunked_array([pa.array(np.arange(13, dtype=np.int64)), a.slice(1), pa.array([None, 4], type=pa.int64()),
                           a.slice(3, 5)])
    col = ingest.import_arrow(ctx, ch)
    back = ingest.column_to_arrow(ctx, col)
    assert back.to_pylist() == ch.to_pylist() and back.null_count == ch.null_count
    col.free()
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


@pytest.mark.gpu
def test_read_parquet_nullable_columns_across_row_groups(ctx, tmp_path):
    import pyarrow as pa
    import pyarrow.parquet as pq
    rng = np.random.default_rng(17)
    n = 70_001
    mask = rng.random(n) < 0.2
    t = pa.table({"a": pa.array(rng.integers(-1 << 40, 1 << 40, n), mask=mask),
                  "b": pa.array(rng.integers(0, 1 << 31, n).astype(np.int32), mask=rng.random(n) < 0.05),
                  "c": pa.array(rng.integers(0, 1 << 32, n).astype(np.uint32))})
    path = tmp_path / "nullable.parquet"
    pq.write_table(t, path, row_group_size=8192)
    cols = ingest.read_parquet(ctx, str(path))
    for name in ("a", "b", "c"):
        back = ingest.column_to_arrow(ctx, cols[name])
        assert back.to_pylist() == t.column(name).to_pylist(), name
        cols[name].free()


@pytest.mark.gpu
def test_q3_from_parquet_files(ctx, orc, tmp_path):
    """customer / orders / lineitem written as Parquet (segments as strings,
    dates as date32, several row groups), read by the C++ reader (strings
    to sorted dictionary codes, dates packed on the device), then the Q3
    chain: every group equals the oracle's on the original tables."""
    import pyarrow as pa
    import pyarrow.parquet as pq
    from pimdal_amd import datagen
    t = datagen.q3_tables(400_000, seed=77)

    def days(packed):
        p = packed.astype(np.int64)
        y, m, d = (p >> 9) + 1900, (p >> 5) & 15, p & 31
        ym = (y - 1970) * 12 + (m - 1)
        return (np.array(ym, dtype="datetime64[M]").astype("datetime64[D]") - np.datetime64("1970-01-01", "D")
                ).astype(np.int64) + d - 1

    segs = np.array(datagen.Q3_SEGMENTS)
    files = {
        "customer": pa.table({"c_custkey": t["c_custkey"], "c_mktsegment": pa.array(segs[t["c_mktsegment"]])}),
        "orders": pa.table({"o_orderkey": t["o_orderkey"], "o_custkey": t["o_custkey"],
                            "o_orderdate": pa.array(days(t["o_orderdate"]).astype(np.int32), type=pa.date32()),
                            "o_shippriority": t["o_shippriority"]}),
        "lineitem": pa.table({"l_orderkey": t["l_orderkey"], "l_extendedprice": t["l_extendedprice"],
                              "l_discount": t["l_discount"],
                              "l_shipdate": pa.array(days(t["l_shipdate"]).astype(np.int32), type=pa.date32())}),
    }
    d = {}
    for name, tab in files.items():
        path = tmp_path / f"{name}.parquet"
        pq.write_table(tab, path, row_group_size=50_000)
        d.update(ingest.read_parquet(ctx, str(path)))
    assert d["c_mktsegment"].dictionary == sorted(datagen.Q3_SEGMENTS)
    assert np.array_equal(d["o_orderdate"].numpy(), t["o_orderdate"])
    seg = d["c_mktsegment"].dictionary.index("BUILDING")
    date = ops.date_pack_str("1995-03-15")
    exp, g_exp = orc.q3(t, datagen.Q3_SEGMENTS.index("BUILDING"), date)
    got, g = ops.tpch_q3(ctx, d, seg, date, capacity=max(g_exp, 1))
    assert g == g_exp and g > 0
    for f in ("key", "revenue", "orderdate", "shippriority"):
        assert np.array_equal(got[f], exp[f]), f
