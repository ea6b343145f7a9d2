"""bench.py's contract checks that need no GPU: the multi-GPU refusal and the
algorithmic byte counts behind the roofline figure."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.parametrize("dim,rbytes,expect", [
    # read x v c type d p m                 write F dc/dt v c
    (3, 4, (24 + 12 + 24 + 4 + 4 + 4 + 4) + (12 + 24 + 12 + 24)),
    (3, 8, (24 + 24 + 48 + 4 + 8 + 8 + 8) + (24 + 48 + 24 + 48)),
    (2, 4, (16 + 8 + 12 + 4 + 4 + 4 + 4) + (8 + 12 + 8 + 12)),
])
def test_pair_pass_algorithmic_bytes(dim, rbytes, expect):
    """k_pairs in a step reads x (double), v, c, type, density, pressure, mass
    and writes the force, dc/dt and the kicked v and c: once per particle."""
    import bench
    assert bench.kernel_bytes("forces", dim, rbytes) == expect


def test_density_algorithmic_bytes():
    import bench
    # x (double) + mass read, density + pressure written
    assert bench.kernel_bytes("density", 3, 4) == 24 + 4 + 8
