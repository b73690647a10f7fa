"""Operator calls over device buffers (the `*_dev` forms of the C ABI).

These mirror PIMDAL's operator set — select, sort, hash join, sort-merge
join, hash and sort aggregation, the fused Q1 filter + group-by — with
device-resident inputs and outputs.  Each call issues HIP kernels through
libpimdal_gpu.so; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call
from .device import Context, DeviceArray, pimdal_type


def _read_count(ctx: Context, d: DeviceArray) -> int:
    ctx.sync()
    return int(d.numpy()[0])


def select(ctx: Context, key: DeviceArray, payload: DeviceArray | None, lo: int, hi: int,
           want_payload: bool = True, want_rowid: bool = True, sync: bool = True):
    """Rows with lo <= key < hi, in row order: (payload, rowid, count).

    reference/db_operators/select.cpp:77-112; pimdal/select/sel.c:59-128.
    """
    n = key.n
    out_p = ctx.empty(n, np.uint32) if want_payload else None
    out_r = ctx.empty(n, np.uint32) if want_rowid else None
    cnt = ctx.empty(1, np.uint64)
    call("pimdal_select_dev", ctx.handle, pimdal_type(key.dtype), key.vp,
         payload.vp if payload is not None else None, n, int(lo), int(hi),
         out_p.vp if out_p is not None else None, out_r.vp if out_r is not None else None, cnt.vp)
    if not sync:
        return out_p, out_r, cnt
    return out_p, out_r, _read_count(ctx, cnt)


def sort_pairs(ctx: Context, keys: DeviceArray, vals: DeviceArray | None, begin_bit: int = -1, end_bit: int = -1,
               out_keys: DeviceArray | None = None, out_vals: DeviceArray | None = None):
    """Stable LSD radix sort of (uint32 key, payload).

    reference/db_operators/sort.cpp:30-70; pimdal/sort/sort_keyval.c:142-179.
    """
    n = keys.n
    ok = out_keys if out_keys is not None else ctx.empty(n, np.uint32)
    vw = 0 if vals is None else vals.dtype.itemsize
    ov = None
    if vals is not None:
        ov = out_vals if out_vals is not None else ctx.empty(n, vals.dtype)
    call("pimdal_sort_pairs_dev", ctx.handle, keys.vp, vals.vp if vals is not None else None, vw, n, ok.vp,
         ov.vp if ov is not None else None, begin_bit, end_bit)
    return ok, ov
