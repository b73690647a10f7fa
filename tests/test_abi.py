"""The C-ABI library loads on the CPU and exports every entry point that
include/*.h declares; the Python binding declares the same set."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


# header -> the library that exports its entry points
LIBS = {"pimdal_parquet.h": "libpimdal_parquet.so"}


def declared(lib="libpimdal_gpu.so"):
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        if LIBS.get(h.name, "libpimdal_gpu.so") != lib:
            continue
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(pimdal_\w+)\s*\(", text, flags=re.M))
    return names


def test_header_declares_entry_points():
    d = declared()
    assert {"pimdal_ctx_create", "pimdal_select", "pimdal_sort_pairs"} <= d


@pytest.mark.parametrize("lib_name", ["libpimdal_gpu.so", "libpimdal_parquet.so"])
def test_library_exports_every_declared_symbol(lib_name):
    import ctypes
    lib_path = ROOT / "pimdal_amd" / lib_name
    if not lib_path.exists():
        pytest.fail(f"{lib_name} is not built (run __graft_entry__.build())")
    ctypes.CDLL(str(ROOT / "pimdal_amd" / "libpimdal_gpu.so"))
    lib = ctypes.CDLL(str(lib_path))
    names = declared(lib_name)
    assert names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from pimdal_amd import _lib
    assert declared() == set(_lib.SIGNATURES)
    assert declared("libpimdal_parquet.so") == set(_lib.PARQUET_SIGNATURES)


def test_load_fails_loudly_without_library(tmp_path, monkeypatch):
    from pimdal_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()
