"""Context and device buffers over the C ABI (numpy in, numpy out)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call

_NP_TYPE = {
    np.dtype(np.uint8): _lib.PIMDAL_U8,
    np.dtype(np.int32): _lib.PIMDAL_I32,
    np.dtype(np.uint32): _lib.PIMDAL_U32,
    np.dtype(np.int64): _lib.PIMDAL_I64,
    np.dtype(np.uint64): _lib.PIMDAL_U64,
}


def pimdal_type(dtype) -> int:
    dt = np.dtype(dtype)
    if dt not in _NP_TYPE:
        raise TypeError(f"unsupported dtype {dt}")
    return _NP_TYPE[dt]


class DeviceArray:
    """A device allocation from the context pool, typed like a 1-D numpy array."""

    def __init__(self, ctx: "Context", n: int, dtype, ptr: int | None = None, owned: bool = True):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        if ptr is None:
            p = C.c_void_p()
            call("pimdal_dev_alloc", ctx.handle, max(self.nbytes, 1), C.byref(p))
            ptr = p.value
        self.ptr = int(ptr)
        self.owned = owned

    def free(self) -> None:
        if self.owned and self.ptr:
            call("pimdal_dev_free", self.ctx.handle, C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            if self.owned and self.ptr and self.ctx.handle:
                _lib.load().pimdal_dev_free(self.ctx.handle, C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass

    def numpy(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else int(n)
        out = np.empty(n, dtype=self.dtype)
        if n:
            call("pimdal_memcpy_d2h", self.ctx.handle, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                 n * self.dtype.itemsize)
        return out

    def view(self, offset: int, n: int) -> "DeviceArray":
        """A non-owning slice [offset, offset + n)."""
        return DeviceArray(self.ctx, n, self.dtype, self.ptr + offset * self.dtype.itemsize, owned=False)

    @property
    def vp(self) -> C.c_void_p:
        return C.c_void_p(self.ptr)


class Context:
    """One GPU, one stream, one HBM budget (pimdal_ctx)."""

    def __init__(self, device: int = 0, budget_bytes: int = 0):
        h = C.c_void_p()
        call("pimdal_ctx_create", device, budget_bytes, C.byref(h))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            call("pimdal_ctx_destroy", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- memory ----
    def empty(self, n: int, dtype) -> DeviceArray:
        return DeviceArray(self, n, dtype)

    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.size, a.dtype)
        if a.size:
            call("pimdal_memcpy_h2d", self.handle, C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes)
        return d

    def sync(self) -> None:
        call("pimdal_ctx_sync", self.handle)

    def mem_stats(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        call("pimdal_ctx_mem_stats", self.handle, *[C.byref(x) for x in v])
        return dict(zip(("live", "reserved", "peak", "budget"), (x.value for x in v)))

    def set_budget(self, b: int) -> None:
        call("pimdal_ctx_set_budget", self.handle, b)

    # ---- profiling ----
    def set_profiling(self, on: bool) -> None:
        call("pimdal_ctx_set_profiling", self.handle, 1 if on else 0)

    def kernel_stat(self, name: str) -> tuple[int, float]:
        n = C.c_uint64()
        ms = C.c_double()
        call("pimdal_ctx_kernel_stat", self.handle, name.encode(), C.byref(n), C.byref(ms))
        return n.value, ms.value

    def reset_stats(self) -> None:
        call("pimdal_ctx_reset_stats", self.handle)

    # ---- tuning ----
    def set_tune(self, key: str, value: int) -> None:
        call("pimdal_ctx_set_tune", self.handle, key.encode(), int(value))

    def clear_tune(self, key: str) -> None:
        """Forget a knob; the library's built-in default applies again."""
        call("pimdal_ctx_clear_tune", self.handle, key.encode())

    def autotune(self, op: str, n: int, cache_path: str | None = None) -> None:
        """Pick the fastest knob values for `op` at size n (pimdal_autotune)."""
        call("pimdal_autotune", self.handle, op.encode(), int(n), cache_path.encode() if cache_path else None)

    def load_tuning(self, cache_path: str) -> None:
        call("pimdal_ctx_load_tuning", self.handle, cache_path.encode())

    def get_tune(self, key: str) -> int:
        v = C.c_int64()
        call("pimdal_ctx_get_tune", self.handle, key.encode(), C.byref(v))
        return v.value
