"""bench.py's command line without a GPU: more than one GPU is refused
with a message, and the traffic figure comes from the committed profile."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


def test_profiled_traffic_is_read_from_the_profile():
    """the committed counter profile holds at least the algorithmic bytes
    of a two-step launch at 4096^2 (18 x 4 B per node) and less than twice"""
    sys.path.insert(0, ROOT)
    import bench
    t = bench.profiled_traffic("k_fused_tb_fp32_ldc4096_hbm_counters.md")
    assert 18 * 4 * 4096 * 4096 <= t < 2 * 18 * 4 * 4096 * 4096
