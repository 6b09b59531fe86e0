"""The device Newton-Krylov engine against the oracle: GMRES on dense systems
with the cases of the reference's solver tests (test/integration/
newton_krylov.jl:10-38), and step_newton! (solver.jl:437-486) on the device
for the Landau and the LB map, against the oracle's Newton step from the same
markers.  Bars: the GMRES iterate to 1e-12 of the oracle's; the Newton step's
state to 1e-9 (both converge on the root of the same implicit system, each
through its own finite-difference products), invariants to roundoff.
"""
import numpy as np
import pytest

import collop
import oracle as O
from conftest import relerr, uniform_bp

pytestmark = pytest.mark.gpu


def test_gmres_dense_matches_oracle_and_numpy(gpu_ctx):
    rng = np.random.default_rng(7)
    n = 60
    A = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, k, r = collop.gmres_dense_gpu(gpu_ctx, A, b, rtol=1e-12)
    xo, ko, ro = O.gmres_dense(A, b, rtol=1e-12, mem=60)
    print(f"dense n=60: device {k} products |r| {r:.2e}; oracle {ko} products |r| {ro:.2e}; "
          f"|x - x_oracle| {np.abs(x - xo).max():.1e}")
    assert k == ko and np.abs(x - xo).max() <= 1e-12
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    assert abs(r - np.linalg.norm(A @ x - b)) <= 1e-10 * np.linalg.norm(b)
    x2, k2, r2 = collop.gmres_dense_gpu(gpu_ctx, A, b, rtol=1e-2)
    assert k2 < k and r2 <= 1e-2 * np.linalg.norm(b)
    assert abs(np.linalg.norm(A @ x2 - b) - r2) <= 1e-6 * r2
    assert collop.gmres_dense_gpu(gpu_ctx, A, b, rtol=1e-14, itmax=3)[1] == 3
    x0, k0, r0 = collop.gmres_dense_gpu(gpu_ctx, A, np.zeros(n), rtol=1e-8)
    assert (k0, r0) == (0, 0.0) and not x0.any()
    with pytest.raises(collop.CollopError):
        collop.gmres_dense_gpu(gpu_ctx, A, b, rtol=1e-8, mem=65)


@pytest.mark.parametrize("model", ["landau", "lb"])
def test_newton_step_matches_oracle(gpu_ctx, model):
    if model == "landau":
        p = collop.config("gaussian_n4096", N_PARTICLES=1024, DT=0.005, solver="newton")
        P = O.default_params(p.DT, m=5, nthreads=0, solver=2)
    else:
        p = collop.config("lb_bimodal_n40000", N_PARTICLES=800, DT=0.001, solver="newton")
        P = O.default_params(p.DT, m=8, nthreads=0, solver=2, collision_model=1, nu=1.0)
    v, w = collop.initial_markers(p)
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    v_ref, fc_ref, st_ref = m.time_step(P, v, w, fc, S0)
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    sim = collop.Simulation(ws, v, w)
    st = sim.run(p.step_params(), 1)[0]
    vg = sim.markers()
    dmax = float(np.max(np.abs(vg - v_ref)))
    print(f"{model}: oracle evals {st_ref.iters} newton {st_ref.n_restart} exit {st_ref.exit_reason}; device evals "
          f"{st['iters']} newton {st['n_restart']} exit {st['exit_reason']} res {st['residual']:.2e}; "
          f"max|dv| {dmax:.2e}")
    assert st["exit_reason"] == 0 and st_ref.exit_reason == 0
    assert st["n_restart"] >= 1 and abs(st["iters"] - st_ref.iters) <= 4
    assert dmax <= 1e-9
    (p1, p2), E0 = O.momentum_energy(v, w)
    assert abs(st["momentum"][0] - p1) <= 1e-13 and abs(st["momentum"][1] - p2) <= 1e-13
    assert abs(st["energy"] - E0) <= 1e-12 * E0 and st["entropy"] >= S0
    assert relerr(sim.coeffs(), fc_ref) <= 1e-9


def test_newton_param_errors(gpu_ctx):
    p = collop.config("gaussian_n4096", N_PARTICLES=600, solver="newton")
    v, w = collop.initial_markers(p)
    ws = collop.Workspace(uniform_bp(-6, 6, 16), uniform_bp(-6, 6, 16), ctx=gpu_ctx)
    sim = collop.Simulation(ws, v, w)
    bad = [p.step_params(precision=collop.PREC_F32C)]
    for field, val in (("nk_krylov_max", 0), ("nk_krylov_max", 65), ("nk_eta_max", 1.0), ("nk_fd_h", -1.0)):
        sp = p.step_params()
        setattr(sp, field, val)
        bad.append(sp)
    for sp in bad:
        with pytest.raises(collop.CollopError):
            sim.run(sp, 1)
