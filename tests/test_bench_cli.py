"""bench.py's contract on the CPU side: more than one GPU is refused with a
message (this build measures one GPU), and the committed counter profile
the roofline block reads names the kernel and carries calibrated traffic."""

import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout.strip() == ""


def test_counter_profile_is_consistent():
    sys.path.insert(0, ROOT)
    import bench
    with open(bench.TRAFFIC_PROFILE) as f:
        prof = json.load(f)
    assert prof["kernel"].startswith("sdk::k_render<false, ")
    want = prof["fetch_kib"] * 1024 * prof["fetch_factor"] + prof["write_kib"] * 1024 * prof["write_factor"]
    assert abs(prof["traffic_bytes"] - want) < 1e-4 * want   # factors are stored to 3 decimals
    assert 0.0 < prof["valu_lane_utilization"] < 1.0 and 0.0 < prof["valu_busy"] <= 1.0
    assert bench.measured_traffic(prof["kernel"]) == prof["traffic_bytes"]
    assert bench.measured_traffic("some other kernel") is None
