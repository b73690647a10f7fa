"""Build libpimdal_gpu.so for gfx950 with hipcc (in-tree, no JIT cache).

Each ``csrc/*.hip`` file is compiled to an object under ``_build/`` in
parallel, then linked into ``pimdal_amd/libpimdal_gpu.so``.  All HIP code is
compiled with packed-FP32 instructions disabled.  ``csrc/parquet_reader.cpp``
(host C++ over the Arrow / Parquet C++ libraries the pyarrow wheel bundles)
becomes ``pimdal_amd/libpimdal_parquet.so``, linked to libpimdal_gpu.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpimdal_gpu.so"
PARQUET_LIB = PKG / "libpimdal_parquet.so"
ARCH = os.environ.get("PIMDAL_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{PKG.parent / 'include'}",
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every csrc/*.hip for gfx950 and link the shared library."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError("no HIP sources found")
    jobs = jobs or min(8, len(sources), os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", flush=True)
    build_parquet(verbose)
    return LIB


def build_parquet(verbose: bool = True) -> Path | None:
    """libpimdal_parquet.so against pyarrow's Arrow / Parquet C++ (C++20, as
    Arrow 25's headers need); skipped when pyarrow is not importable."""
    try:
        import pyarrow
    except ImportError:
        if verbose:
            print("libpimdal_parquet.so skipped: pyarrow not importable", flush=True)
        return None
    pa_dir = Path(pyarrow.__file__).resolve().parent
    gxx = shutil.which("g++") or "g++"
    tmp = PARQUET_LIB.with_suffix(".so.tmp")
    cmd = [gxx, "-std=c++20", "-O2", "-fPIC", "-shared", "-Wall", f"-I{PKG.parent / 'include'}",
           f"-I{pa_dir / 'include'}", str(CSRC / "parquet_reader.cpp"), "-o", str(tmp),
           f"-L{pa_dir}", "-l:libparquet.so.2500", "-l:libarrow.so.2500", f"-Wl,-rpath,{pa_dir}",
           f"-L{PKG}", "-l:libpimdal_gpu.so", "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"libpimdal_parquet.so build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, PARQUET_LIB)
    if verbose:
        print(f"built {PARQUET_LIB}", flush=True)
    return PARQUET_LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
