"""Host-side checks of bench.py's contract: the committed counter summaries it
reads the roofline's measured traffic from parse, and a request for more than
one GPU ends with a message instead of a measurement."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("kernel", ["k_force", "k_gather", "k_node"])
@pytest.mark.parametrize("mixed", [False, True])
def test_traffic_read_from_profiles(kernel, mixed):
    t = bench.read_traffic(kernel, mixed)
    assert t is not None, f"profiles/ has no parsable traffic line for {kernel} mixed={mixed}"
    # measured HBM bytes of one launch lie within a small factor of the algorithmic bytes
    alg = bench.kernel_bytes(kernel, 1_749_570, 102_324_536, mixed)
    assert 0.5 * alg < t < 4.0 * alg


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one MI355X" in r.stderr
    assert r.stdout.strip() == ""
