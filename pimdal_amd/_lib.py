"""ctypes binding of libpimdal_gpu.so (the C ABI in include/pimdal_gpu.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpimdal_gpu.so"

c_ctx_p = C.c_void_p
c_col_p = C.c_void_p
u64 = C.c_uint64
i64 = C.c_int64
u32 = C.c_uint32
i32 = C.c_int32
vp = C.c_void_p
dbl = C.c_double

PIMDAL_U8, PIMDAL_I32, PIMDAL_U32, PIMDAL_I64, PIMDAL_U64, PIMDAL_I128 = 1, 2, 3, 4, 5, 6

ERRORS = {
    -1: "PIMDAL_E_INVALID",
    -2: "PIMDAL_E_UNSUPPORTED",
    -3: "PIMDAL_E_NOMEM",
    -4: "PIMDAL_E_HIP",
    -5: "PIMDAL_E_CAPACITY",
    -6: "PIMDAL_E_INTERNAL",
}


class PimdalError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


# name -> argtypes; every function returns int (0 = ok) unless listed in _RESTYPE
SIGNATURES: dict[str, list] = {
    "pimdal_last_error": [],
    "pimdal_version": [],
    "pimdal_ctx_create": [i32, u64, C.POINTER(c_ctx_p)],
    "pimdal_ctx_destroy": [c_ctx_p],
    "pimdal_ctx_sync": [c_ctx_p],
    "pimdal_ctx_stream": [c_ctx_p, C.POINTER(vp)],
    "pimdal_ctx_mem_stats": [c_ctx_p, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)],
    "pimdal_ctx_set_budget": [c_ctx_p, u64],
    "pimdal_ctx_set_profiling": [c_ctx_p, C.c_int],
    "pimdal_ctx_kernel_stat": [c_ctx_p, C.c_char_p, C.POINTER(u64), C.POINTER(dbl)],
    "pimdal_ctx_reset_stats": [c_ctx_p],
    "pimdal_ctx_set_tune": [c_ctx_p, C.c_char_p, i64],
    "pimdal_ctx_clear_tune": [c_ctx_p, C.c_char_p],
    "pimdal_ctx_get_tune": [c_ctx_p, C.c_char_p, C.POINTER(i64)],
    "pimdal_autotune": [c_ctx_p, C.c_char_p, u64, C.c_char_p],
    "pimdal_ctx_load_tuning": [c_ctx_p, C.c_char_p],
    "pimdal_dev_alloc": [c_ctx_p, u64, C.POINTER(vp)],
    "pimdal_dev_free": [c_ctx_p, vp],
    "pimdal_dev_copy": [c_ctx_p, vp, vp, u64],
    "pimdal_memcpy_h2d": [c_ctx_p, vp, vp, u64],
    "pimdal_memcpy_d2h": [c_ctx_p, vp, vp, u64],
    "pimdal_col_alloc": [c_ctx_p, C.c_int, u64, C.POINTER(c_col_p)],
    "pimdal_col_from_host": [c_ctx_p, C.c_int, vp, u64, C.POINTER(c_col_p)],
    "pimdal_col_wrap": [c_ctx_p, C.c_int, vp, u64, C.POINTER(c_col_p)],
    "pimdal_col_to_host": [c_ctx_p, c_col_p, vp, u64],
    "pimdal_col_info": [c_col_p, C.POINTER(C.c_int), C.POINTER(u64), C.POINTER(vp)],
    "pimdal_col_free": [c_ctx_p, c_col_p],
    "pimdal_col_validity": [c_col_p, C.POINTER(vp), C.POINTER(u64)],
    "pimdal_col_set_validity": [c_ctx_p, c_col_p, vp],
    "pimdal_col_import": [c_ctx_p, vp, vp, C.POINTER(c_col_p)],
    "pimdal_col_import_chunks": [c_ctx_p, vp, u32, vp, C.POINTER(c_col_p)],
    "pimdal_date_pack": [i32, i32, i32, C.POINTER(i32)],
    "pimdal_date_pack_str": [C.c_char_p, C.POINTER(i32)],
    "pimdal_date32_pack_dev": [c_ctx_p, vp, u64, vp],
    "pimdal_col_export": [c_ctx_p, c_col_p, vp, vp],
    "pimdal_select": [c_ctx_p, c_col_p, c_col_p, i64, i64, C.POINTER(c_col_p), C.POINTER(c_col_p)],
    "pimdal_select_dev": [c_ctx_p, C.c_int, vp, vp, u64, i64, i64, vp, vp, vp],
    "pimdal_select_dev_noalloc": [c_ctx_p, C.c_int, vp, vp, u64, i64, i64, vp, vp, vp],
    "pimdal_sort_pairs": [c_ctx_p, c_col_p, c_col_p, C.POINTER(c_col_p), C.POINTER(c_col_p)],
    "pimdal_sort_pairs_dev": [c_ctx_p, vp, vp, u32, u64, vp, vp, C.c_int, C.c_int],
    "pimdal_hash_join": [c_ctx_p, c_col_p, c_col_p, c_col_p, c_col_p, C.POINTER(c_col_p), C.POINTER(c_col_p)],
    "pimdal_hash_join_dev": [c_ctx_p, vp, vp, u64, vp, vp, u64, vp, vp, u64, C.POINTER(u64)],
    "pimdal_sort_merge_join": [c_ctx_p, c_col_p, c_col_p, c_col_p, c_col_p, C.POINTER(c_col_p), C.POINTER(c_col_p)],
    "pimdal_sort_merge_join_dev": [c_ctx_p, vp, vp, u64, vp, vp, u64, vp, vp, u64, C.POINTER(u64)],
    "pimdal_bloom_build_dev": [c_ctx_p, vp, u64, u32, u32, vp],
    "pimdal_bloom_probe_dev": [c_ctx_p, vp, vp, u64, u32, u32, vp, vp, vp, vp],
    "pimdal_bloom_semi_join_dev": [c_ctx_p, vp, u64, vp, vp, u64, u32, u32, vp, vp, vp],
    "pimdal_hash_aggr_dev": [c_ctx_p, vp, vp, u64, vp, vp, vp, u64, C.POINTER(u64)],
    "pimdal_sort_aggr_dev": [c_ctx_p, vp, vp, u64, vp, vp, vp, u64, C.POINTER(u64)],
    "pimdal_hash_aggr": [c_ctx_p, c_col_p, c_col_p, C.POINTER(c_col_p), C.POINTER(c_col_p), C.POINTER(c_col_p)],
    "pimdal_sort_aggr": [c_ctx_p, c_col_p, c_col_p, C.POINTER(c_col_p), C.POINTER(c_col_p), C.POINTER(c_col_p)],
    "pimdal_filter_groupby_q1_dev": [c_ctx_p, vp, vp, vp, vp, vp, vp, vp, vp, u64, i32, vp, u64, C.POINTER(u64)],
    "pimdal_q6_dev": [c_ctx_p, vp, vp, vp, vp, u64, i32, i32, C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                      C.POINTER(u64)],
    "pimdal_filter_groupby_q1_dev_async": [c_ctx_p, vp, vp, vp, vp, vp, vp, vp, vp, u64, i32, vp, u64, vp, vp],
    "pimdal_q6_dev_async": [c_ctx_p, vp, vp, vp, vp, u64, i32, i32, C.c_int64, C.c_int64, C.c_int64, vp, vp],
    "pimdal_tpch_q3_dev": [c_ctx_p, vp, u32, i32, vp, u64, C.POINTER(u64)],
    "pimdal_tpch_q4_dev": [c_ctx_p, vp, i32, i32, vp, vp, u64, C.POINTER(u64)],
    "pimdal_tpch_q5_dev": [c_ctx_p, vp, u32, i32, i32, vp, vp, u64, C.POINTER(u64)],
    "pimdal_plan_begin": [c_ctx_p, C.POINTER(vp)],
    "pimdal_plan_end": [vp],
    "pimdal_plan_launch": [vp],
    "pimdal_plan_destroy": [vp],
    "pimdal_server_create": [c_ctx_p, vp, C.c_int, u64, C.POINTER(vp)],
    "pimdal_server_run": [vp, vp, u64, C.c_int, vp, vp],
    "pimdal_server_destroy": [vp],
    "pimdal_gen_dim_keys": [c_ctx_p, u64, u64, vp],
    "pimdal_gen_u32": [c_ctx_p, u64, u32, u64, vp],
    "pimdal_gen_range": [c_ctx_p, C.c_int, u64, u32, u64, i64, i64, vp],
    "pimdal_gen_perm": [c_ctx_p, u64, u32, u64, u64, u64, vp],
    "pimdal_perm_apply": [c_ctx_p, u64, u32, u64, u64, i64, vp, vp],
    "pimdal_gen_iota": [c_ctx_p, u64, u32, vp],
    "pimdal_gen_zipf": [c_ctx_p, u64, u32, u64, dbl, u64, vp],
    "pimdal_fmix32": [c_ctx_p, u64, vp, vp],
    "pimdal_gen_orderkey": [c_ctx_p, u64, u64, u64, vp],
}

_RESTYPE = {"pimdal_last_error": C.c_char_p, "pimdal_parquet_last_error": C.c_char_p}

# libpimdal_parquet.so (include/pimdal_parquet.h)
PARQUET_LIB_PATH = Path(__file__).resolve().parent / "libpimdal_parquet.so"
PARQUET_SIGNATURES: dict[str, list] = {
    "pimdal_parquet_last_error": [],
    "pimdal_parquet_shape": [C.c_char_p, C.POINTER(u32), C.POINTER(u64)],
    "pimdal_parquet_column_name": [C.c_char_p, u32, C.c_char_p, u64],
    "pimdal_parquet_read_column": [c_ctx_p, C.c_char_p, C.c_char_p, i32, u32, C.POINTER(c_col_p), C.c_char_p, u64,
                                   C.POINTER(u32)],
}
PARQUET_PACK_DATES, PARQUET_DICT_CODES = 1, 2
_parquet: C.CDLL | None = None


def load_parquet() -> C.CDLL:
    """Load libpimdal_parquet.so (raises if it is missing)."""
    global _parquet
    if _parquet is not None:
        return _parquet
    load()
    if not PARQUET_LIB_PATH.exists():
        raise RuntimeError(f"{PARQUET_LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(PARQUET_LIB_PATH))
    for name, args in PARQUET_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _parquet = lib
    return lib


def call_parquet(name: str, *args) -> None:
    rc = getattr(load_parquet(), name)(*args)
    if rc != 0:
        raise PimdalError(rc, (load_parquet().pimdal_parquet_last_error() or b"").decode())

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the built library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() or python pimdal_amd/build.py"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def last_error() -> str:
    return (load().pimdal_last_error() or b"").decode()


def check(rc: int) -> None:
    if rc != 0:
        raise PimdalError(rc, last_error())


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
