"""Arrow and Parquet columns in and out of device memory.

`read_parquet` mirrors the TPC-H table reader of the reference
(pimdal/main/tpc_h/reader/read_table.cpp:18-35, parquet_to_table: one
Parquet file, the selected columns by index) through libpimdal_parquet.so,
which reads each column with Parquet C++ and imports it through the Arrow
C Data Interface into one device column (pimdal_col_import_chunks: the
chunks, one per row group, concatenated).
date32 columns become the reference's packed dates on the device
(read_table.cpp:8-16; pimdal_date32_pack_dev), so they compare directly
with `operators.date_pack_str` literals, unless `pack_dates` is False.
Nullable columns keep their validity bitmap on the device, and the column
forms below (select, joins, aggregations) give NULLs their SQL meaning.
`to_arrow` / `column_to_arrow` are the way back: host Arrow arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call
from .device import Context, DeviceArray, pimdal_type


class _ArrowArray(C.Structure):
    # struct ArrowArray of the Arrow C Data Interface (include/pimdal_arrow_c.h)
    _fields_ = [("length", C.c_int64), ("null_count", C.c_int64), ("offset", C.c_int64),
                ("n_buffers", C.c_int64), ("n_children", C.c_int64), ("buffers", C.c_void_p),
                ("children", C.c_void_p), ("dictionary", C.c_void_p), ("release", C.c_void_p),
                ("private_data", C.c_void_p)]


class _ArrowSchema(C.Structure):
    # struct ArrowSchema of the Arrow C Data Interface
    _fields_ = [("format", C.c_char_p), ("name", C.c_char_p), ("metadata", C.c_char_p), ("flags", C.c_int64),
                ("n_children", C.c_int64), ("children", C.c_void_p), ("dictionary", C.c_void_p),
                ("release", C.c_void_p), ("private_data", C.c_void_p)]


_RELEASE_ARRAY = C.CFUNCTYPE(None, C.POINTER(_ArrowArray))
_RELEASE_SCHEMA = C.CFUNCTYPE(None, C.POINTER(_ArrowSchema))

_DTYPE = {
    _lib.PIMDAL_U8: np.dtype(np.uint8),
    _lib.PIMDAL_I32: np.dtype(np.int32),
    _lib.PIMDAL_U32: np.dtype(np.uint32),
    _lib.PIMDAL_I64: np.dtype(np.int64),
    _lib.PIMDAL_U64: np.dtype(np.uint64),
}


def _release(s) -> None:
    if s.release:
        fn = _RELEASE_ARRAY if isinstance(s, _ArrowArray) else _RELEASE_SCHEMA
        fn(s.release)(C.byref(s))


class ColumnArray(DeviceArray):
    """A device column owned by a pimdal_col handle (released with pimdal_col_free)."""

    def __init__(self, ctx: Context, col: C.c_void_p):
        t, n, p = C.c_int(), C.c_uint64(), C.c_void_p()
        call("pimdal_col_info", col, C.byref(t), C.byref(n), C.byref(p))
        if t.value not in _DTYPE:
            call("pimdal_col_free", ctx.handle, col)
            raise TypeError(f"column type {t.value} has no numpy dtype")
        super().__init__(ctx, n.value, _DTYPE[t.value], ptr=p.value or 0, owned=False)
        self.col = col

    def free(self) -> None:
        if self.col:
            call("pimdal_col_free", self.ctx.handle, self.col)
        self.col = None
        self.ptr = 0

    def __del__(self):
        try:
            if self.col and self.ctx.handle:
                _lib.load().pimdal_col_free(self.ctx.handle, self.col)
                self.col = None
        except Exception:
            pass


def import_arrow(ctx: Context, arr) -> ColumnArray:
    """One device column from a pyarrow Array or ChunkedArray (chunks in order)."""
    import pyarrow as pa
    chunks = arr.chunks if isinstance(arr, pa.ChunkedArray) else [arr]
    k = len(chunks)
    arrays = (_ArrowArray * max(k, 1))()
    schema = _ArrowSchema()
    exported = 0
    try:
        arr.type._export_to_c(C.addressof(schema))
        for i, ch in enumerate(chunks):
            ch._export_to_c(C.addressof(arrays[i]))
            exported += 1
        ptrs = (C.c_void_p * max(k, 1))(*[C.addressof(arrays[i]) for i in range(k)])
        col = C.c_void_p()
        call("pimdal_col_import_chunks", ctx.handle, ptrs, k, C.addressof(schema), C.byref(col))
    finally:
        for i in range(exported):
            _release(arrays[i])
        _release(schema)
    return ColumnArray(ctx, col)


def to_arrow(ctx: Context, a: DeviceArray):
    """A host pyarrow Array holding the values of device column `a`."""
    import pyarrow as pa
    col = C.c_void_p()
    call("pimdal_col_wrap", ctx.handle, pimdal_type(a.dtype), a.vp, a.n, C.byref(col))
    out_a, out_s = _ArrowArray(), _ArrowSchema()
    try:
        call("pimdal_col_export", ctx.handle, col, C.addressof(out_a), C.addressof(out_s))
    finally:
        call("pimdal_col_free", ctx.handle, col)
    return pa.Array._import_from_c(C.addressof(out_a), C.addressof(out_s))


def read_parquet(ctx: Context, path, columns=None, pack_dates: bool = True, dict_codes: bool = True) -> dict:
    """Columns of a Parquet file as device columns: {name: ColumnArray},
    read by libpimdal_parquet.so (pimdal_parquet_read_column: C++ over
    Arrow / Parquet C++, every row group into one device column, validity
    kept).

    `columns` selects by index (the `sel_col` of read_table.cpp:18-35) or by
    name; None reads every column.  date32 columns are packed on the device
    unless `pack_dates` is False (then they hold days since 1970-01-01).
    String columns become sorted dictionary codes (uint8 up to 256 values)
    with the dictionary in the column's `dictionary` attribute, unless
    `dict_codes` is False (then they are refused).
    """
    n_cols = C.c_uint32()
    n_rows = C.c_uint64()
    _lib.call_parquet("pimdal_parquet_shape", str(path).encode(), C.byref(n_cols), C.byref(n_rows))
    sel = list(range(n_cols.value)) if columns is None else list(columns)
    flags = (_lib.PARQUET_PACK_DATES if pack_dates else 0) | (_lib.PARQUET_DICT_CODES if dict_codes else 0)
    out = {}
    try:
        for c in sel:
            if isinstance(c, (int, np.integer)):
                buf = C.create_string_buffer(4096)
                _lib.call_parquet("pimdal_parquet_column_name", str(path).encode(), int(c), buf, 4096)
                name, index = buf.value.decode(), int(c)
            else:
                name, index = str(c), -1
            col = C.c_void_p()
            nd = C.c_uint32()
            dbuf = C.create_string_buffer(1 << 20)
            _lib.call_parquet("pimdal_parquet_read_column", ctx.handle, str(path).encode(), name.encode(), index,
                              flags, C.byref(col), dbuf, 1 << 20, C.byref(nd))
            arr = ColumnArray(ctx, col)
            arr.dictionary = [v.decode() for v in dbuf.raw.split(b"\0")[:nd.value]] if nd.value else None
            out[name] = arr
    except Exception:
        for a in out.values():
            a.free()
        raise
    return out


def column_to_arrow(ctx: Context, col: "ColumnArray"):
    """A host pyarrow Array of a library column, with its validity."""
    import pyarrow as pa
    out_a, out_s = _ArrowArray(), _ArrowSchema()
    call("pimdal_col_export", ctx.handle, col.col, C.addressof(out_a), C.addressof(out_s))
    return pa.Array._import_from_c(C.addressof(out_a), C.addressof(out_s))


def _col(ctx: Context, handle: C.c_void_p) -> "ColumnArray | None":
    return ColumnArray(ctx, handle) if handle.value else None


def col_select(ctx: Context, key: ColumnArray, payload: ColumnArray | None, lo: int, hi: int):
    """pimdal_select over library columns: (payload column, row-id column).
    A NULL key satisfies no predicate; payload validity follows the rows."""
    op, orow = C.c_void_p(), C.c_void_p()
    call("pimdal_select", ctx.handle, key.col, payload.col if payload is not None else None, int(lo), int(hi),
         C.byref(op), C.byref(orow))
    return _col(ctx, op), _col(ctx, orow)


def col_join(ctx: Context, build_key: ColumnArray, build_payload, probe_key: ColumnArray, probe_payload,
             sort_merge: bool = False):
    """pimdal_hash_join / pimdal_sort_merge_join over library columns:
    (probe payload column, build payload column).  NULL keys join nothing."""
    op, ob = C.c_void_p(), C.c_void_p()
    call("pimdal_sort_merge_join" if sort_merge else "pimdal_hash_join", ctx.handle, build_key.col,
         build_payload.col if build_payload is not None else None, probe_key.col,
         probe_payload.col if probe_payload is not None else None, C.byref(op), C.byref(ob))
    return _col(ctx, op), _col(ctx, ob)


def col_aggr(ctx: Context, key: ColumnArray, val: ColumnArray, sort: bool = False):
    """pimdal_hash_aggr / pimdal_sort_aggr over library columns: (key, sum,
    count) columns.  With NULLs: the NULL-key rows form the last group, SUM
    skips NULL values, COUNT counts non-NULL values."""
    ok, os_, oc = C.c_void_p(), C.c_void_p(), C.c_void_p()
    call("pimdal_sort_aggr" if sort else "pimdal_hash_aggr", ctx.handle, key.col, val.col, C.byref(ok),
         C.byref(os_), C.byref(oc))
    return _col(ctx, ok), _col(ctx, os_), _col(ctx, oc)
