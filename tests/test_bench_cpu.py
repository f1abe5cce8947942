"""bench.py's host-side contract (no GPU): it refuses more than one GPU with a message, and
reads the measured per-iteration HBM traffic from the committed D32 profile."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_measured_traffic_comes_from_the_profile():
    for prec, alg in (("float64", 104.0), ("float32", 52.0)):
        t = bench.measured_traffic(prec)
        assert t is not None
        # the fused iteration moves about the textbook 13 words: float64 a little under them (the
        # G1+V pass keeps its cell intermediate on chip), float32 more (the cell intermediates and
        # composite rows in float32 weigh more against the halved edge words), but well under 1.5x
        assert 0.9 * alg * 40986624 < t < 1.5 * alg * 40986624
