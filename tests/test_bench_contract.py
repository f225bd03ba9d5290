"""bench.py's command-line contract, on the CPU: one GPU only (more exits with a message and
status 2 before touching a device), and the metric string the JSON line carries."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                          text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr
    assert proc.stdout == ""


def test_metric_string_and_flags():
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.pop(0)
    assert bench.METRIC == "GRG matmul edge·RHS/s + HBM GB/s frac; served p50/p99 vs plain, 1 GPU"
    assert bench.K_LADDER == [1, 2, 4, 8, 16, 32, 64, 128, 256] and sum(bench.K_LADDER) == 511
    # algorithmic bytes of one traversal: 4(E+V) + s*K*(E+V+2*C_in+2*C_out)
    assert bench.algorithmic_bytes(10, 20, 2, 4, 3, 5) == 4 * 30 + 4 * 2 * (30 + 6 + 10)
