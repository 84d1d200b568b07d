"""bench.py's host-side contract: one GPU only, a learned form's warm-up that
reaches the steady state, and counters taken over the timed steps only."""
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


def test_learned_warmup_reaches_steady_state():
    """A learned form's warm-up covers the substeps after which the network
    has passed its first check (order 2: one substep per step)."""
    import bench
    for form in ("learned", "learned-cubic", "learned-p8"):
        P = bench.learned_config(form)["period"]
        assert bench.learned_warmup(form, 5) >= 6 * P
        assert bench.learned_warmup(form, 1000) == 1000


def test_stats_delta_subtracts_counters_only():
    import bench
    before = {k: 1 for k in bench.COUNTERS}
    before.update(last_check_err=0.5)
    after = {k: 4 for k in bench.COUNTERS}
    after.update(last_check_err=0.25)
    d = bench.stats_delta(after, before)
    assert all(d[k] == 3 for k in bench.COUNTERS)
    assert d["last_check_err"] == 0.25
