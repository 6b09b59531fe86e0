"""The device-resident defect-correction solve (step_defect!, solver.jl:569-634)
against the oracle's restatement, and against the Anderson solve of the same
step: both must return a root of the same implicit system.
"""
import numpy as np
import pytest

import collop
import oracle as O
from conftest import relerr, uniform_bp

pytestmark = pytest.mark.gpu


def _oracle_residual(m, P, v_new, v, w, S0):
    return float(np.linalg.norm(m.picard_map(P, v_new, v, w, S0) - v_new))


@pytest.mark.parametrize("ext", [(0, 0.0), (3, 0.04)], ids=["reference", "outer_aa+near"])
@pytest.mark.parametrize("ic,N,ne,dt,strict", [("gaussian", 1024, 16, 0.005, True),
                                               ("bimodal", 2048, 32, 0.02, True),
                                               ("gaussian", 1024, 16, 0.05, False)])
def test_defect_step_matches_oracle(gpu_ctx, ic, N, ne, dt, strict, ext):
    sig = (1.0, 1.0) if ic == "gaussian" else (np.sqrt(0.5), np.sqrt(0.5))
    v, w = collop.markers(ic, N, 42, sigma1=sig[0], sigma2=sig[1], peak=2.0)
    bp = uniform_bp(-6, 6, ne)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(dt, m=5, nthreads=0, solver=1, dc_outer_m=ext[0], dc_near_theta=ext[1])
    v_ref, fc_ref, st_ref = m.time_step(P, v, w, fc, S0)

    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    sim = collop.Simulation(ws, v, w)
    prm = collop.config("gaussian_n4096", DT=dt, solver="defect", dc_outer_m=ext[0],
                        dc_near_theta=ext[1]).step_params()
    st = sim.run(prm, 1)[0]
    v_gpu = sim.markers()
    res = _oracle_residual(m, P, v_gpu, v, w, S0)
    tol = max(1e-12 * np.linalg.norm(v_gpu), 1e-10)
    print(f"{ic} N={N} dt={dt} ext={ext}: oracle outer {st_ref.iters} inner {st_ref.inner} exit {st_ref.exit_reason}; "
          f"gpu outer {st['iters']} inner {st['inner_maps']} exit {st['exit_reason']}; "
          f"max|dv| {np.max(np.abs(v_gpu - v_ref)):.2e}; oracle-map residual {res:.2e} (tol {tol:.2e})")
    assert st["exit_reason"] == 0 and st_ref.exit_reason == 0
    assert res <= 20 * tol
    (p1, p2), E0 = O.momentum_energy(v, w)
    assert abs(st["energy"] - E0) <= 1e-12 * E0
    assert abs(st["momentum"][0] - p1) <= 1e-13 and abs(st["momentum"][1] - p2) <= 1e-13
    assert st["entropy"] >= S0
    if strict:
        assert abs(st["iters"] - st_ref.iters) <= 1
        assert abs(st["inner_maps"] - st_ref.inner) <= max(3, st_ref.inner // 5)
        assert np.max(np.abs(v_gpu - v_ref)) <= 1e-9
        assert relerr(sim.coeffs(), fc_ref) <= 1e-9


def test_defect_and_anderson_reach_the_same_root(gpu_ctx):
    p = collop.config("bimodal_n65536", N_PARTICLES=8192, DT=0.005)
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, ctx=gpu_ctx)
    a = collop.Simulation(ws, v, w)
    sa = a.run(p.step_params(), 1)[0]
    va = a.markers()
    b = collop.Simulation(ws, v, w)
    sb = b.run(p.step_params(solver=collop._lib.SOLVER_DEFECT), 1)[0]
    vb = b.markers()
    print(f"anderson maps {sa['iters']}; defect outer {sb['iters']} inner {sb['inner_maps']}; "
          f"max|dv| {np.max(np.abs(va - vb)):.2e}")
    assert sa["exit_reason"] == 0 and sb["exit_reason"] == 0
    assert sb["iters"] < sa["iters"]
    assert np.max(np.abs(va - vb)) <= 1e-9


@pytest.mark.parametrize("ext", [(0, 0.0), (3, 0.04)], ids=["reference", "outer_aa+near"])
def test_defect_conservation_bimodal_n65536(gpu_ctx, ext):
    p = collop.config("bimodal_n65536", solver="defect", dc_outer_m=ext[0], dc_near_theta=ext[1])
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, ctx=gpu_ctx)
    sim = collop.Simulation(ws, v, w)
    E0 = 0.5 * np.sum(w * (v[:, 0] ** 2 + v[:, 1] ** 2))
    P_prev = (w[:, None] * v).sum(axis=0)
    S_prev = collop.compute_r_and_entropy_gpu(ws, sim.coeffs())[1]
    E_prev = E0
    for s in sim.run(p.step_params(), 3):
        print({k: s[k] for k in ("iters", "inner_maps", "exit_reason", "residual", "entropy", "seconds")})
        assert s["exit_reason"] == 0
        assert abs(s["energy"] - E_prev) / E0 <= 1e-10
        assert np.all(np.abs(np.array(s["momentum"]) - P_prev) <= 1e-10)
        assert s["entropy"] >= S_prev - 1e-12
        S_prev, E_prev, P_prev = s["entropy"], s["energy"], np.array(s["momentum"])


def test_defect_param_errors(gpu_ctx):
    bp = uniform_bp(-6, 6, 16)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    v, w = collop.markers("gaussian", 64, 1)
    sim = collop.Simulation(ws, v, w)
    bad = collop.config("gaussian_n4096", solver="defect").step_params()
    bad.dc_eta = 1.5
    with pytest.raises(collop.CollopError) as ei:
        sim.run(bad, 1)
    assert ei.value.status == collop._lib.ERR_INVALID
    bad = collop.config("gaussian_n4096", solver="defect", dc_outer_m=9).step_params()
    with pytest.raises(collop.CollopError):
        sim.run(bad, 1)
    bad = collop.config("gaussian_n4096").step_params()
    bad.solver = 7
    with pytest.raises(collop.CollopError):
        sim.run(bad, 1)
    bad = collop.config("gaussian_n4096", solver="defect", dc_eta_f64=1.5).step_params()
    with pytest.raises(collop.CollopError):
        sim.run(bad, 1)


def test_defect_eta_f64_matches_oracle(gpu_ctx):
    """dc_eta_f64 sets the inner tolerance of the FP64 outer iterations; with
    FP64 maps throughout that is every outer iteration, i.e. the oracle's
    step_defect! with dc_eta = dc_eta_f64."""
    v, w = collop.markers("gaussian", 1024, 42)
    bp = uniform_bp(-6, 6, 16)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(0.005, m=5, nthreads=0, solver=1, dc_eta=0.002)
    v_ref, _, st_ref = m.time_step(P, v, w, fc, S0)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    sim = collop.Simulation(ws, v, w)
    st = sim.run(collop.config("gaussian_n4096", DT=0.005, solver="defect", dc_eta_f64=0.002).step_params(), 1)[0]
    dv = float(np.max(np.abs(sim.markers() - v_ref)))
    print(f"eta_f64 0.002: oracle outer {st_ref.iters} inner {st_ref.inner}; gpu outer {st['iters']} "
          f"inner {st['inner_maps']}; max|dv| {dv:.2e}")
    assert st["exit_reason"] == 0 and st_ref.exit_reason == 0
    assert abs(st["iters"] - st_ref.iters) <= 1
    assert abs(st["inner_maps"] - st_ref.inner) <= max(3, st_ref.inner // 5)
    assert dv <= 1e-9


@pytest.mark.parametrize("dt", [0.005, 0.02])
def test_defect_f32c_outer_maps_reach_the_fp64_root(gpu_ctx, dt):
    """The chosen path (FP32-compensated outer maps until the FP32 floor, FP64
    outer maps reusing the FP32c metric, outer Anderson, near field) lands on
    the root of the all-FP64 solve."""
    p = collop.defect_config("bimodal_n65536", N_PARTICLES=8192, DT=dt)
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, ctx=gpu_ctx)
    a = collop.Simulation(ws, v, w)
    sa = a.run(p.step_params(precision=collop.PREC_F64), 1)[0]
    va = a.markers()
    b = collop.Simulation(ws, v, w)
    sb = b.run(p.step_params(precision=collop.PREC_F32C), 1)[0]
    vb = b.markers()
    print(f"dt={dt}: fp64 outer {sa['iters']} inner {sa['inner_maps']}; f32c outer {sb['iters']} "
          f"(f32 {sb['iters_f32']}) inner {sb['inner_maps']}; max|dv| {np.max(np.abs(va - vb)):.2e}")
    assert sa["exit_reason"] == 0 and sb["exit_reason"] == 0
    assert sb["iters_f32"] >= 1
    assert np.max(np.abs(va - vb)) <= 1e-9
