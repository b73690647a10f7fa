"""Serving loop bindings: resident tables, a query stream, two run modes.

`Server.run(queries, mode)` calls pimdal_server_run (C++ scheduler in
libpimdal_gpu.so); this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call
from .device import Context

SELECT, JOIN, GROUPBY = 1, 2, 3
SEQUENTIAL, BATCHED = 0, 1


class Tables(C.Structure):
    _fields_ = [("orderkey", C.c_void_p), ("quantity", C.c_void_p), ("price", C.c_void_p), ("discount", C.c_void_p),
                ("tax", C.c_void_p), ("date", C.c_void_p), ("flag1", C.c_void_p), ("flag2", C.c_void_p),
                ("n_fact", C.c_uint64), ("dim_key", C.c_void_p), ("dim_date", C.c_void_p),
                ("dim_category", C.c_void_p), ("n_dim", C.c_uint64)]


QUERY_DTYPE = np.dtype([("kind", np.int32), ("tenant", np.uint32), ("offset", np.uint64), ("rows", np.uint64),
                        ("param", np.int64), ("arrival_us", np.float64)])
RESULT_DTYPE = np.dtype([("count", np.uint64), ("checksum", np.uint64), ("start_us", np.float64),
                         ("end_us", np.float64), ("latency_us", np.float64)])


class Stats(C.Structure):
    _fields_ = [("makespan_us", C.c_double), ("qps", C.c_double), ("p50_us", C.c_double), ("p99_us", C.c_double),
                ("mean_us", C.c_double), ("batches", C.c_uint64), ("max_inflight", C.c_uint64),
                ("heap_bytes", C.c_uint64), ("heap_peak_bytes", C.c_uint64), ("tenants", C.c_uint64)]


class Server:
    def __init__(self, ctx: Context, fact: dict, dim: dict, slots: int = 4, heap_bytes: int = 0):
        self.ctx = ctx
        self._keep = (fact, dim)
        t = Tables(fact["orderkey"].ptr, fact["quantity"].ptr, fact["price"].ptr, fact["discount"].ptr,
                   fact["tax"].ptr, fact["date"].ptr, fact["flag1"].ptr, fact["flag2"].ptr, fact["date"].n,
                   dim["key"].ptr, dim["date"].ptr, dim["category"].ptr, dim["key"].n)
        h = C.c_void_p()
        call("pimdal_server_create", ctx.handle, C.byref(t), slots, heap_bytes, C.byref(h))
        self.handle = h

    def run(self, queries: np.ndarray, mode: int):
        q = np.ascontiguousarray(queries, dtype=QUERY_DTYPE)
        res = np.zeros(q.size, dtype=RESULT_DTYPE)
        st = Stats()
        call("pimdal_server_run", self.handle, q.ctypes.data_as(C.c_void_p), q.size, mode,
             res.ctypes.data_as(C.c_void_p), C.byref(st))
        return res, {f: getattr(st, f) for f, _ in Stats._fields_}

    def close(self):
        if self.handle:
            call("pimdal_server_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_queries(n: int, n_fact: int, rate_qps: float, seed: int = 0x5EED0005, min_log2: float = 16,
                max_log2: float = 26, mix=(0.5, 0.3, 0.2), tenants: int = 1) -> np.ndarray:
    """Open-loop Poisson arrivals; 50/30/20 select/join/group-by; slice sizes
    log-uniform in [2^min_log2, 2^max_log2]; selectivity s ~ U(0.01, 0.5);
    each query's tenant uniform over `tenants`."""
    rng = np.random.default_rng(seed)
    q = np.zeros(n, dtype=QUERY_DTYPE)
    if tenants > 1:
        q["tenant"] = rng.integers(0, tenants, size=n)
    q["kind"] = rng.choice([SELECT, JOIN, GROUPBY], size=n, p=list(mix))
    rows = np.floor(2.0 ** rng.uniform(min_log2, max_log2, size=n)).astype(np.uint64)
    rows = np.minimum(rows, n_fact).astype(np.uint64)
    q["rows"] = rows
    q["offset"] = (rng.random(n) * (n_fact - rows + 1)).astype(np.uint64)
    s = rng.uniform(0.01, 0.5, size=n)
    param = np.where(q["kind"] == SELECT, 90000 + s * (10500000 - 90000), s * 2557)
    q["param"] = np.floor(param).astype(np.int64)
    gaps = rng.exponential(1e6 / rate_qps, size=n) if rate_qps > 0 else np.zeros(n)
    q["arrival_us"] = np.cumsum(gaps) - gaps[0]
    return q
