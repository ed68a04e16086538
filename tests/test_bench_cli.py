"""bench.py's command line and byte accounting, on the CPU (no GPU calls):
more than one GPU is refused with a message, every workload has smoothing
defaults, bad smoothing counts are rejected, and the algorithmic bytes of a
residual sweep follow the stated formula."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

WORKLOADS = ["poisson128", "adjoint96", "stokes64", "q2_64", "q2ada", "ilu128"]


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_every_workload_has_smoothing_defaults():
    assert sorted(bench.SMOOTH) == sorted(WORKLOADS)
    for fp64, mixed in bench.SMOOTH.values():
        bench.parse_smooth(fp64)
        bench.parse_smooth(mixed)
    # the Q1 Poisson FP64 path is the reference's algorithm, V(2,2)
    assert bench.parse_smooth(bench.SMOOTH["poisson128"][0]) == (2, 2)


@pytest.mark.parametrize("text", ["3", "3,2,1", "a,b", "-1,2", "40,2"])
def test_bad_smoothing_counts_are_rejected(text):
    with pytest.raises((SystemExit, ValueError)):
        bench.parse_smooth(text)


def test_residual_bytes():
    """stored values + column bytes + 12 B per slice + 2 B per row + x, b, r"""
    lay = dict(stored=1000, column_bytes=300, slices=4)
    n = 200
    assert bench.sell_bytes(n, lay, 8, 8) == 1000 * 8 + 300 + 4 * 12 + 2 * n + 3 * 8 * n
    assert bench.sell_bytes(n, lay, 2, 4) == 1000 * 2 + 300 + 4 * 12 + 2 * n + 3 * 4 * n
