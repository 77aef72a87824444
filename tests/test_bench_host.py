"""Host-side checks of the benchmark script (no GPU): the one-GPU rule and the committed
profile summaries its roofline object reads."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_profile_summaries_parse_for_the_benchmark_frame():
    traffic = bench.read_traffic("k_primary", 1920, 1080)
    assert traffic is not None and 50e6 < traffic < 200e6  # planes 66 MB + queue tens of MB
    c = bench.read_counters("k_primary", 1920, 1080)
    assert set(c) == {"valu_busy", "valu_lane_utilisation", "waves_per_cu"}
    assert 0.0 < c["valu_busy"] <= 1.0 and 0.0 < c["valu_lane_utilisation"] <= 1.0
    assert 0.0 < c["waves_per_cu"] <= 32.0
    # another frame size has no committed summary
    assert bench.read_traffic("k_primary", 320, 180) is None and bench.read_counters("k_primary", 320, 180) == {}
