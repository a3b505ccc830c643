"""bench.py's contract pieces that run without a GPU: the one-GPU guard and
the PMC traffic figures it reads from the committed profile summaries."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("name, algorithmic", [
    ("stem_fit_kernel-stem_fit_64x64x4096x8.md", 8 * 134217728),
    ("ray_program_kernel-plain_1M.md", 36 * 1048576 + 128 * 128 * 4),
    ("stem_forward_kernel-stem_fwd_64x64x4096.md", 12 * 16777216),
    ("stem_backproject_kernel-backproject_64x64x4096.md", 4 * 16777216),
])
def test_profiled_traffic_is_read_and_at_least_algorithmic(name, algorithmic):
    import bench

    t = bench._profiled_traffic(name)
    assert isinstance(t, int)
    assert algorithmic <= t <= 5 * algorithmic
