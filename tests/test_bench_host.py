"""bench.py's host logic without a GPU: the metric's constraint over a
sequence of step statistics, and the refusal of more than one GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _st(E, P=(0.0, 0.0), S=1.0, exit_reason=0, res=1e-10, handover=0):
    return {"energy": E, "momentum": P, "entropy": S, "exit_reason": exit_reason, "iters": 10, "iters_f32": 2,
            "inner_maps": 0, "r0": 0.3, "residual": res, "n_clusters": 0, "cl_fail": 0, "handover": handover}


def test_constraint_met_only_when_every_step_converges_and_conserves():
    E0 = 2.0
    good = [_st(2.0 + 1e-11, S=1.1), _st(2.0, S=1.2, handover=3), _st(2.0 - 1e-11, S=1.3)]
    c = bench.constraint_summary(good, E0, 2.0, (0.0, 0.0), 1.0)
    assert c["constraint_met"] and c["all_converged"] and c["handed_over_steps"] == 1
    assert c["max_dE_rel"] == pytest.approx(1e-11 / E0) and c["min_dS"] == pytest.approx(0.1)
    # one stagnated step fails it even with the energy conserved
    c = bench.constraint_summary(good[:2] + [_st(2.0, S=1.3, exit_reason=1, res=3e-4)], E0, 2.0, (0.0, 0.0), 1.0)
    assert not c["constraint_met"] and c["stagnated_steps"] == 1 and c["max_exit_residual"] == 3e-4
    # energy drift above 1e-10 per step (relative to E0), momentum drift above 1e-10
    c = bench.constraint_summary([_st(2.0 + 3e-10, S=1.1)], E0, 2.0, (0.0, 0.0), 1.0)
    assert not c["constraint_met"] and c["max_dE_rel"] == pytest.approx(1.5e-10)
    c = bench.constraint_summary([_st(2.0, P=(2e-10, 0.0), S=1.1)], E0, 2.0, (0.0, 0.0), 1.0)
    assert not c["constraint_met"] and c["max_dP"] == pytest.approx(2e-10)
    # the entropy falls on the second step
    c = bench.constraint_summary([_st(2.0, S=1.1), _st(2.0, S=1.05)], E0, 2.0, (0.0, 0.0), 1.0)
    assert not c["constraint_met"] and c["min_dS"] == pytest.approx(-0.05)
    # a step that hit the iteration cap
    c = bench.constraint_summary([_st(2.0, S=1.1, exit_reason=2)], E0, 2.0, (0.0, 0.0), 1.0)
    assert not c["constraint_met"] and c["max_iter_steps"] == 1
    # drifts are per step: two steps of 0.8e-10 each meet it
    c = bench.constraint_summary([_st(2.0 + 1.6e-10, S=1.1), _st(2.0 + 3.2e-10, S=1.2)], E0, 2.0, np.zeros(2), 1.0)
    assert c["constraint_met"]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
