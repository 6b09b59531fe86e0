"""The fused small-kernel maps against the launch-per-stage maps they replace:
the frozen-metric inner maps of defect correction (scatter with the pending
Anderson update, fused quadrature, G_eff / metric product / near pairs inside
the epilogue) and the Lenard-Bernstein Anderson maps (entropy quadrature
inside the gather launch; reduction, lambda, drift multipliers and v_dot
inside the epilogue).  Both forms run the
same arithmetic in the same order, so every step must agree bit for bit:
markers, iteration counts and residuals.
"""
import numpy as np
import pytest

import collop

pytestmark = pytest.mark.gpu


def _run(ws, v, w, prm, steps, fused):
    collop.set_fused_maps(ws, fused)
    try:
        sim = collop.Simulation(ws, v, w)
        stats = sim.run(prm, steps)
        return sim.markers(), stats
    finally:
        collop.set_fused_maps(ws, True)


def _same(a, b):
    (va, sa), (vb, sb) = a, b
    assert np.array_equal(va, vb), f"max |dv| {np.max(np.abs(va - vb)):.3e}"
    for x, y in zip(sa, sb):
        for key in ("iters", "exit_reason", "residual", "inner_maps", "energy", "entropy"):
            assert x[key] == y[key], (key, x[key], y[key])


@pytest.mark.parametrize("gonzalez", [True, False])
def test_lb_fused_maps_bitwise(gpu_ctx, gonzalez):
    p = collop.config("lb_bimodal_n40000", N_PARTICLES=16000, DT=0.001)
    v, w = collop.initial_markers(p)
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    prm = p.step_params()
    prm.use_gonzalez = 1 if gonzalez else 0
    fused = _run(ws, v, w, prm, 3, True)
    staged = _run(ws, v, w, prm, 3, False)
    print("LB iters", [s["iters"] for s in fused[1]], "gonzalez", gonzalez)
    assert all(s["iters"] > 2 for s in fused[1])  # Anderson updates between fused maps
    _same(fused, staged)


@pytest.mark.parametrize("outer_m,theta,fp32", [(0, 0.0, False), (3, 0.02, True)], ids=["reference", "chosen"])
def test_defect_fused_frozen_maps_bitwise(gpu_ctx, outer_m, theta, fp32):
    p = collop.config("bimodal_n65536", N_PARTICLES=4096, solver="defect", gpu_fp32=fp32, dc_outer_m=outer_m,
                      dc_near_theta=theta)
    v, w = collop.initial_markers(p)
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    prm = p.step_params()
    fused = _run(ws, v, w, prm, 2, True)
    staged = _run(ws, v, w, prm, 2, False)
    print("defect inner maps", [s["inner_maps"] for s in fused[1]])
    assert all(s["inner_maps"] > 2 for s in fused[1])
    _same(fused, staged)


def test_fused_maps_switch_errors(gpu_ctx):
    p = collop.config("lb_bimodal_n40000", N_PARTICLES=64)
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    with pytest.raises(collop.CollopError):
        ws.ctx.check(ws.ctx.lib.collop_set_fused_maps(ws.ctx.h, 2))
