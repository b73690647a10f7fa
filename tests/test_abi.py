"""The C-ABI library loads on the CPU and exports every entry point that
include/*.h declares; the Python binding declares the same set."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(pimdal_\w+)\s*\(", text, flags=re.M))
    return names


def test_header_declares_entry_points():
    d = declared()
    assert {"pimdal_ctx_create", "pimdal_select", "pimdal_sort_pairs"} <= d


def test_library_exports_every_declared_symbol():
    import ctypes
    lib_path = ROOT / "pimdal_amd" / "libpimdal_gpu.so"
    if not lib_path.exists():
        pytest.fail("libpimdal_gpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(lib_path))
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from pimdal_amd import _lib
    assert declared() == set(_lib.SIGNATURES)


def test_load_fails_loudly_without_library(tmp_path, monkeypatch):
    from pimdal_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()
