"""CPU checks of bench.py's contract that need no GPU: it measures one GPU and
refuses more, and its roofline traffic figure is the committed profile's."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    assert "one GPU" in out.stderr and out.stdout == ""


def test_roofline_traffic_is_the_committed_profile():
    sys.path.insert(0, ROOT)
    import bench
    with open(os.path.join(ROOT, "profiles", "ntt_256x2^20_hbm_traffic.json")) as f:
        want = json.load(f)["hbm_bytes_per_call"]
    assert bench.measured_traffic(20, 256) == want
    assert bench.measured_traffic(21, 7) is None  # no profile for that shape
