"""Writes tests/golden/gen_digests.json: SHA-256 of small generated columns.

Run: python -m tests.golden.make_digests   (from the repository root)
The digests pin the generators so that data cannot drift between the host
restatement and the device generators without a test noticing.
"""
import hashlib
import json
from pathlib import Path

import numpy as np


def _d(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def compute(orc):
    n = 4096
    k, p = orc.select_workload(n)
    sk, _ = orc.sort_workload(n, zipf=True)
    rk, rp, jk, jp = orc.join_workload(1 << 10, n, zipf=True)
    t = orc.fact_table(n)
    out = {
        "select_key": _d(k), "select_pay": _d(p), "sort_zipf_key": _d(sk),
        "join_build_key": _d(rk), "join_build_pay": _d(rp), "join_probe_key": _d(jk), "join_probe_pay": _d(jp),
    }
    for c, v in sorted(t.items()):
        out["fact_" + c] = _d(v)
    return out


if __name__ == "__main__":
    import sys
    root = Path(__file__).resolve().parents[2]
    sys.path.insert(0, str(root))
    from oracle import orc
    path = Path(__file__).resolve().parent / "gen_digests.json"
    path.write_text(json.dumps(compute(orc), indent=1, sort_keys=True) + "\n")
    print("wrote", path)
