"""bench.py's host-side contract: one GPU only, and the kernel labels it
attaches to its roofline entries follow the library's kernel choice."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("form,n,name", [
    ("fp64", 2048, "k_accel<FP64, IPT=1> (tile)"),
    ("mixed", 16383, "k_accel<MIXED_NR, IPT=1> (tile)"),
    ("fp64", 16384, "k_accel_sym<FP64, K=4> (symmetric)"),
    ("mixed", 131071, "k_accel_sym<MIXED_NR, K=4> (symmetric)"),
    ("fp64", 262144, "k_accel_sym<FP64, K=8> (symmetric)"),
    ("mixed", 262144, "k_accel_sym<MIXED_NR, K=6> (symmetric)"),
])
def test_force_kernel_names(form, n, name):
    import bench
    assert bench.force_kernel_name(form, n) == name
