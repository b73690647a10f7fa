"""bench.py's host side: the CPU baseline's cases and host description, and
the refusal of more than one GPU (no GPU needed)."""
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_cpulist_parsing():
    assert bench._cpulist("0-3,8,10-11\n") == {0, 1, 2, 3, 8, 10, 11}
    assert bench._cpulist("") == set()


def test_host_info_fields():
    info = bench.host_info()
    assert set(info) >= {"cpu_model", "numa_nodes", "affinity_cpus", "online_cpus", "cgroup_cpu_quota"}
    assert info["affinity_cpus"] >= 1


def test_cpu_cases_name_gpu_lines():
    """Every case carries a GPU operator line's name, a positive row count and
    a sample description; the names are distinct."""
    names = [c[0] for c in bench._cpu_cases()]
    assert len(names) == len(set(names))
    gpu_lines = {"sort_2^28", "sort_zipf_2^28", "select_2^27", "hash_join", "hash_join_zipf", "smj_join",
                 "smj_join_zipf", "bloom_semi_join_2^16x2^27", "bloom_semi_join_2^20x2^27", "q1_flags",
                 "q1_orderkey", "q6", "q3_sf10", "q4_sf10", "q5_sf10", "date32_pack_6e7"}
    gpu_lines |= {f"{w}_aggr_2^27_{s}" for w, s in (("hash", "k50"), ("hash", "k4000"), ("hash", "k2^24"),
                                                     ("sort", "k25"), ("sort", "k2^24"))}
    assert set(names) == gpu_lines
    for name, sample, rows, make, run in bench._cpu_cases():
        assert rows > 0 and sample


def test_aggr_workload_recipe():
    from oracle import orc
    k, v = orc.aggr_workload(10_000, 1, 50)
    assert k.dtype == np.uint32 and v.dtype == np.int64
    assert k.min() >= 1 and k.max() <= 50 and v.min() >= 1 and v.max() <= 10
    k2, v2 = orc.aggr_workload(10_000, 1, 50)
    assert np.array_equal(k, k2) and np.array_equal(v, v2)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2", "--no-cpu"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr
