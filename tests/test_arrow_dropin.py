"""The header-only Arrow C++ adapter (include/pimdal_arrow.hpp) used the way a
caller of reference/db_operators would: compiled here, run on the GPU, and its
outputs compared with the CPU restatement."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.util import same_multiset

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "arrow_dropin" / "dropin_main.cpp"


def build_dropin(out: Path) -> Path:
    import pyarrow
    pa = Path(pyarrow.__file__).parent
    exe = out / "dropin"
    cmd = ["g++", "-std=c++20", "-O2", str(SRC), "-o", str(exe), f"-I{ROOT / 'include'}", f"-I{pa / 'include'}",
           f"-L{pa}", "-l:libarrow.so.2500", f"-Wl,-rpath,{pa}", f"-L{ROOT / 'pimdal_amd'}", "-l:libpimdal_gpu.so",
           f"-Wl,-rpath,{ROOT / 'pimdal_amd'}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_adapter_compiles_and_links(tmp_path):
    if not (ROOT / "pimdal_amd" / "libpimdal_gpu.so").exists():
        pytest.fail("libpimdal_gpu.so not built")
    assert build_dropin(tmp_path).exists()


@pytest.mark.gpu
def test_adapter_results_match_oracle(tmp_path, orc):
    exe = build_dropin(tmp_path)
    res = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    rd = lambda name, dt: np.fromfile(tmp_path / name, dtype=dt)
    n = 1 << 20
    # regenerate the caller's inputs with the same counter-based generators
    key = orc.gen_range(np.int32, 0xD0, 1, n, 1, 50)
    ukey = orc.gen_u32(0xD0, 2, n)
    inner = orc.gen_perm(0xD0, 3, n // 4, n // 4) + np.uint32(1)
    outer = orc.gen_range(np.uint32, 0xD0, 4, n, 1, n // 4)
    op, orow = orc.select(key, key.view(np.uint32), 10, 30)
    assert np.array_equal(rd("select_key", np.int32), op.view(np.int32))
    assert np.array_equal(rd("select_row", np.uint32), orow)
    assert np.array_equal(rd("sort_key", np.uint32), orc.sort_keys(ukey))
    ep, eb = orc.hash_join(inner, inner, outer, outer)
    assert same_multiset((rd("hj_outer", np.uint32), rd("hj_inner", np.uint32)), (ep, eb))
    mp, mb = orc.sort_merge_join(inner, inner, outer, outer)
    assert np.array_equal(rd("smj_outer", np.uint32), mp) and np.array_equal(rd("smj_inner", np.uint32), mb)
    val = (1 + np.arange(n) % 10).astype(np.int64)
    ek, es, ec = orc.hash_aggr(key.view(np.uint32), val)
    gk, gs, gc = rd("agg_key", np.int32), rd("agg_sum", np.int64), rd("agg_cnt", np.int64)
    o = np.argsort(gk)
    assert np.array_equal(gk[o].view(np.uint32), ek) and np.array_equal(gs[o], es) and np.array_equal(gc[o], ec)
