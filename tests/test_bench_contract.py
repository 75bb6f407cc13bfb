"""bench.py contract pieces that need no GPU: the multi-GPU refusal, and the
profiled-counter summary it reads for the headline workload."""
from __future__ import annotations

import json
import os
import subprocess
import sys
import types

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_profiled_counts_parse_the_committed_summary():
    sys.path.insert(0, ROOT)
    import bench

    a = types.SimpleNamespace(ebn0=4.0, osd_order=1)
    got = bench.profiled_counts(a, 1 << 20)
    assert got is not None
    # HBM bytes per launch: at least the algorithmic 577 B per codeword, and no
    # wholesale re-reads; VALU wave-instructions per codeword in a plausible band
    assert 577 * (1 << 20) <= got["traffic_bytes_per_launch"] <= 2 * 577 * (1 << 20)
    assert 100 <= got["valu_insts_per_launch"] / (1 << 20) <= 1000
    # other workloads carry no profiled figures
    assert bench.profiled_counts(a, 1 << 19) is None
