"""bench.py's command-line contract, checked without a GPU."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    p = _bench("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_every_workload_is_offered():
    p = _bench("--help")
    assert p.returncode == 0
    for name in ("controller", "spacecurve-batch", "dense-d32", "plain-d4", "uneven", "wbasis-d8"):
        assert name in p.stdout


def test_level_counts_of_the_bench_match_the_library_statement():
    """The generic workloads' roofline uses bench._lm_counts; for a fused level it
    states what the library's timer states (samples and previous level read once,
    level written once; products plus scan flops)."""
    sys.path.insert(0, str(ROOT))
    import bench

    B, Q, K, S, d = 3, 1, 4, 129, 8
    m = d * d
    by, fl = bench._lm_counts(B, Q, K, S, m, m, 2 * d ** 3)
    assert by["lm_fwd_fused"] == by["lm_fwd_prod"] == 8 * (B * S * m + 2 * B * S * m)
    assert fl["lm_fwd_fused"] == fl["lm_fwd_prod"] + fl["lm_fwd_scan"]
