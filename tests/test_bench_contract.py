"""bench.py contract checks that need no GPU: one GPU only, argument parsing,
and the algorithmic-bytes / traffic bookkeeping."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr or "single MI355X" in p.stderr


def test_measured_traffic_is_only_reported_for_the_profiled_workload():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic(3, (64, 64, 64))
    f = os.path.join(ROOT, "profiles", "k_dg_apply_q3_64cubed_traffic.json")
    assert t == json.load(open(f))["traffic_bytes_per_launch"]
    assert bench.measured_traffic(3, (32, 32, 32)) is None
    assert bench.measured_traffic(2, (64, 64, 64)) is None
