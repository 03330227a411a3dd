"""bench.py on the CPU: the one-GPU rule and the link between the dominant
kernel's launch shape and the traffic figure recorded in profiles/ (without
it the JSON line's roofline.traffic would silently be null)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "measures one GPU" in p.stderr
    assert p.stdout == ""


def test_fused_grid_threads_counts_whole_waves_of_workgroups():
    # 256^3, 3 steps, 2 rows: regions of 58 x 26 outputs, 52-plane chunks
    tiles = -(-256 // 58) * -(-256 // 26) * -(-256 // 52)
    assert bench.fused_grid_threads(256, 3, 2, 52) == -(-tiles // 8) * 8 * 1024


@pytest.mark.parametrize("rows,ichunk", [(2, 52), (4, 26)])
def test_profile_holds_traffic_of_both_finest_level_launch_shapes(rows, ichunk):
    grid = bench.fused_grid_threads(256, 3, rows, ichunk)
    t = bench.measured_traffic(f"k_poly_fused<float,3,{rows}>", 256, ichunk, grid)
    assert t is not None
    # at least the algorithmic bytes (read x, b; write x: 12 B per node)
    assert t >= 12 * 256 ** 3
    # a shape the profile does not hold gives None, not a wrong figure
    assert bench.measured_traffic(f"k_poly_fused<float,3,{rows}>", 256, ichunk + 1, grid) is None
