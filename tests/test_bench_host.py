"""bench.py's host-side pieces on the CPU: the algorithmic byte model, the figures it reads from the
committed profile summaries, the one-GPU refusal, and the FIFO replay of the plain loop's latencies."""
import os
import subprocess
import sys

import pytest

from conftest import REPO

sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_step_bytes_are_the_sum_of_the_kernel_figures():
    """One FCT + alpha stage = alpha + low-order + R- + apply passes, a third of the dt pass, and the
    16 B of U^n the RK combination reads: 324 + 70.67k bytes per node-stage (818.7 at k = 7)."""
    for k in (6.0, 6.998, 7.0, 9.5):
        per = sum(bench.kernel_bytes_per_node(n, k) for n in ("k_alpha", "k_fct_lo", "k_fct_rminus", "k_fct_apply"))
        per += bench.kernel_bytes_per_node("k_compute_dt", k) / 3.0 + 16
        assert abs(per - bench.step_bytes_per_node_stage(k)) < 0.01 * k
    assert abs(bench.step_bytes_per_node_stage(7.0) - 818.69) < 0.01
    # the SGN model's stage pass reads and writes two more components
    assert bench.kernel_bytes_per_node("k_stage_lo", 7.0, True, 5) - bench.kernel_bytes_per_node("k_stage_lo", 7.0, True) \
        == pytest.approx(16 + 16 + 16 * 2 / 3)


def test_profile_figures_are_found_for_the_headline_workload():
    t = bench.measured_traffic("k_fct_lo", "bumps_pd", 16716273)
    assert t is not None and 3.0e9 < t < 1.0e10
    busy, lanes = bench.measured_valu("k_fct_lo", "bumps_pd", 16716273)
    assert 0.0 < busy <= 1.0 and 0.0 < lanes <= 1.0
    assert bench.measured_traffic("k_fct_lo", "bumps_pd", 1000) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_fifo_replay():
    """Requests served one at a time in arrival order: each starts when it arrives or when the one
    before it ends, whichever is later."""
    fin = bench.replay_fifo([0.0, 0.5, 10.0], [2.0, 1.0, 1.0])
    assert fin == pytest.approx([2.0, 3.0, 11.0])
