"""bench.py's contract pieces that need no GPU: more than one GPU is refused
with a message, and the committed traffic summary is read per launch shape."""

from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_traffic_is_read_per_launch_shape():
    sys.path.insert(0, ROOT)
    import bench

    a, b = bench.read_traffic((2, 48)), bench.read_traffic((4, 48))
    assert a is not None and b is not None and a > 0 and b > 0
    # bytes per launch are above the 52 B/cell algorithmic traffic of 64 x 600 x 1200 cells
    assert min(a, b) > 52 * 64 * 600 * 1200
    assert bench.read_traffic((3, 999)) is None
