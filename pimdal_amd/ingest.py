"""Arrow and Parquet columns in and out of device memory.

`read_parquet` mirrors the TPC-H table reader of the reference
(pimdal/main/tpc_h/reader/read_table.cpp:18-35, parquet_to_table: one
Parquet file, the selected columns by index) and imports every column
through the Arrow C Data Interface into one device column
(pimdal_col_import_chunks: the chunks, one per row group, concatenated).
date32 columns become the reference's packed dates on the device
(read_table.cpp:8-16; pimdal_date32_pack_dev), so they compare directly
with `operators.date_pack_str` literals, unless `pack_dates` is False.
Nullable columns are refused (PIMDAL_E_UNSUPPORTED), as the operators take
none.  `to_arrow` is the way back: a host Arrow array of a device column.
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


def read_parquet(ctx: Context, path, columns=None, pack_dates: bool = True) -> dict:
    """Columns of a Parquet file as device columns: {name: ColumnArray}.

    `columns` selects by index (the `sel_col` of read_table.cpp:18-35) or by
    name; None reads every column.  date32 columns are packed in place on
    the device unless `pack_dates` is False (then they hold days since
    1970-01-01).
    """
    import pyarrow as pa
    import pyarrow.parquet as pq
    if columns is not None:
        names = pq.ParquetFile(path).schema_arrow.names
        columns = [names[c] if isinstance(c, (int, np.integer)) else c for c in columns]
    table = pq.read_table(path, columns=columns)
    out = {}
    try:
        for name in table.column_names:
            col = table.column(name)
            out[name] = import_arrow(ctx, col)
            if pack_dates and pa.types.is_date32(col.type):
                from .operators import date32_pack
                date32_pack(ctx, out[name], out[name])
    except Exception:
        for c in out.values():
            c.free()
        raise
    return out
