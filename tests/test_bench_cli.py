"""bench.py's command-line contract that holds without a GPU: more than one GPU is
refused with a message, the canonical byte and operation models of the roofline
figures."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_canonical_bytes_of_the_bubble_stage():
    # per node 32*Ns + (d^2+1)*8 + Naux*8, per face point (d+1)*8 + 8, per element 2d*4
    assert bench.algorithmic_bytes_per_stage(3, 5, 4096, 5, 4) == 163938304
    assert bench.algorithmic_bytes_per_stage(2, 5, 1, 4, 0) == 25 * 168 + 20 * 32 + 16


def test_ec_operation_count():
    # per node, direction and partner (self included): 3d^2+16d+48 (+15 gravity) + Ns(6d+2)
    assert bench.ec_volume_flops_per_stage(3, 5, 4096, 5, True) == 4096 * 125 * 3 * 5 * (123 + 15 + 100)
    assert bench.ec_volume_flops_per_stage(2, 5, 1, 4, False) == 25 * 2 * 5 * (12 + 32 + 48 + 4 * 14)
