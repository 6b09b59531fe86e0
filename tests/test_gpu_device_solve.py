"""The device-resident Anderson solve against the host-controlled one.

With collop_set_device_solve(ctx, 1) (the default) a control kernel per map
reduces the epilogue statistics, runs step_anderson!'s state machine
(solver.jl:300-423: the window's least squares of solver.jl:396-412, the exit,
best-iterate, restart and stagnation tests) and plans the next map, and the
host only launches work; with 0 the host reads every map's statistics and
decides.  The state machine is the same code in both, compiled without
floating-point contraction, so every step must agree bit for bit: markers,
iteration counts, exit reasons and residuals.  The exits covered: convergence
(with and without the exit Picard step), stagnation and the map budget.
"""
import numpy as np
import pytest

import collop

pytestmark = pytest.mark.gpu

KEYS = ("iters", "iters_f32", "n_restart", "exit_reason", "residual", "r0", "energy", "entropy")


def _run(ws, v, w, prm, steps, device):
    collop.set_device_solve(ws, device)
    try:
        sim = collop.Simulation(ws, v, w)
        stats = sim.run(prm, steps)
        return sim.markers(), stats
    finally:
        collop.set_device_solve(ws, True)


def _same(a, b):
    (va, sa), (vb, sb) = a, b
    assert np.array_equal(va, vb), f"max |dv| {np.max(np.abs(va - vb)):.3e}"
    for x, y in zip(sa, sb):
        for key in KEYS:
            assert x[key] == y[key], (key, x[key], y[key])


def _setup(gpu_ctx, N, **kw):
    p = collop.config("bimodal_n65536", N_PARTICLES=N, **kw)
    v, w = collop.initial_markers(p)
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    return p, v, w, ws


@pytest.mark.parametrize("fp32,exit_picard,m", [(False, 1, 5), (True, 1, 5), (False, 0, 2), (True, 0, 3)],
                         ids=["f64", "f32c", "f64-no-exit-step-m2", "f32c-no-exit-step-m3"])
def test_device_solve_bitwise(gpu_ctx, fp32, exit_picard, m):
    p, v, w, ws = _setup(gpu_ctx, 4096, gpu_fp32=fp32)
    prm = p.step_params()
    prm.exit_picard_step = exit_picard
    prm.m_anderson = m
    dev = _run(ws, v, w, prm, 3, True)
    host = _run(ws, v, w, prm, 3, False)
    print("iters", [s["iters"] for s in dev[1]], "f32", [s["iters_f32"] for s in dev[1]],
          "syncs", [s["stat_syncs"] for s in dev[1]])
    assert all(s["exit_reason"] == 0 for s in dev[1])
    assert all(s["iters"] > 3 for s in dev[1])
    if fp32:
        assert all(0 < s["iters_f32"] < s["iters"] for s in dev[1])
    _same(dev, host)
    assert all(s["stat_syncs"] == 0 for s in dev[1])


@pytest.mark.parametrize("gonzalez,logsq", [(0, 0), (1, 1)], ids=["no-gonzalez", "half-log-squared"])
def test_device_solve_map_variants(gpu_ctx, gonzalez, logsq):
    """The map without the Gonzalez correction, and with the half log f^2
    entropy: both controls still agree bit for bit."""
    p, v, w, ws = _setup(gpu_ctx, 4096, gpu_fp32=True)
    prm = p.step_params()
    prm.use_gonzalez = gonzalez
    prm.use_logsq = logsq
    dev = _run(ws, v, w, prm, 2, True)
    host = _run(ws, v, w, prm, 2, False)
    _same(dev, host)


def test_device_solve_stagnation_and_budget(gpu_ctx):
    """Forced stagnation exits and map-budget exits end on the best iterate in both."""
    p, v, w, ws = _setup(gpu_ctx, 4096)
    prm = p.step_params()
    prm.stag_window = 3
    prm.stag_rel_tol = 0.99
    dev = _run(ws, v, w, prm, 2, True)
    host = _run(ws, v, w, prm, 2, False)
    assert all(s["exit_reason"] == 1 for s in dev[1])
    _same(dev, host)
    prm = p.step_params()
    prm.max_iter = 5
    prm.tol = 1e-30
    prm.abs_floor = 1e-30
    dev = _run(ws, v, w, prm, 2, True)
    host = _run(ws, v, w, prm, 2, False)
    assert all(s["exit_reason"] == 2 and s["iters"] == 5 for s in dev[1])
    _same(dev, host)


def test_device_solve_bimodal_n65536(gpu_ctx):
    """Two steps of the bench workload (N = 65536, FP32c maps then FP64 maps)."""
    p, v, w, ws = _setup(gpu_ctx, 65536, gpu_fp32=True)
    prm = p.step_params()
    dev = _run(ws, v, w, prm, 2, True)
    host = _run(ws, v, w, prm, 2, False)
    print("iters", [s["iters"] for s in dev[1]], "f32", [s["iters_f32"] for s in dev[1]])
    _same(dev, host)


def test_device_solve_switch_errors(gpu_ctx):
    p = collop.config("bimodal_n65536", N_PARTICLES=64)
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    with pytest.raises(collop.CollopError):
        ws.ctx.check(ws.ctx.lib.collop_set_device_solve(ws.ctx.h, 2))
