"""bench.py's contract pieces that need no GPU (CPU only).

* More than one GPU is refused with a message and a nonzero status, before any device work.
* The `traffic` figure comes from the committed counter summary: a positive byte count per
  trace launch for every node layout, and the summary names the calibration factors used.
* The guide network's FLOP count per row matches its layer shapes (96 -> 64 -> 64 -> 64 -> 64 -> 48).
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_multi_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr


def test_traffic_comes_from_the_committed_summary():
    import bench
    for layout in ("2", "4", "4q"):
        t = bench.measured_traffic(layout)
        assert isinstance(t, int) and t > 0, (layout, t)
    text = open(os.path.join(ROOT, "profiles", "k_trace_refill_terrain3M_pmc.md")).read()
    assert "FETCH_SIZE" in text and "WRITE_SIZE" in text and "scale factor" in text


def test_mlp_flops_per_row():
    import bench
    layers = [(96, 64), (64, 64), (64, 64), (64, 64), (64, 48)]
    assert bench.MLP_FLOP_PER_ROW == sum(2 * k * n for k, n in layers) == 43008
