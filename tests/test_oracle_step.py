"""Pins the oracle's implicit step (solver.jl:62-423, main.jl:226-307) by the
properties the scheme guarantees: the Gonzalez discrete chain rule holds to
roundoff on every map, a map conserves momentum and energy at the midpoint,
Anderson and plain damped Picard reach the same fixed point, the step
conserves momentum/energy and does not decrease entropy, and the solver's
exits (convergence, stagnation, iteration cap) behave as written.
"""
import numpy as np
import pytest

import collop
import oracle as O
from conftest import uniform_bp


@pytest.fixture(scope="module")
def small():
    p = collop.config("gaussian_n4096", N_PARTICLES=600)
    v, w = collop.initial_markers(p)
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(p.DT, m=5, nthreads=4)
    v1, _ = m.predictor(P, v, w, fc)
    return dict(p=p, v=v, w=w, m=m, fc=fc, S0=S0, P=P, v1=v1)


def test_gonzalez_discrete_chain_rule(small):
    m, v, w, S0 = small["m"], small["v"], small["w"], small["S0"]
    vin = small["v1"]
    Geff, vmid, lam = m.landau_geff(vin, v, w, S0)
    S1 = m.entropy(m.l2_project(vin, w))
    dv = vin - v
    # dbar S . dv = S(v_in) - S0 with dbar S = -w G_eff  (solver.jl:137-149)
    lhs = np.sum(-w[:, None] * Geff * dv)
    assert abs(lhs - (S1 - S0)) < 1e-13 * max(1.0, abs(S1 - S0)) + 1e-15
    assert lam != 0.0
    Gm, _, lam0 = m.landau_geff(vin, v, w, S0, use_gonzalez=False)
    assert lam0 == 0.0
    np.testing.assert_allclose(Gm - lam * dv / w[:, None], Geff, rtol=0, atol=1e-12 * np.abs(Geff).max())
    np.testing.assert_allclose(vmid, 0.5 * (v + vin), rtol=0, atol=0)


def test_picard_map_conserves_at_midpoint(small):
    m, v, w, S0, P = small["m"], small["v"], small["w"], small["S0"], small["P"]
    vin = small["v1"]
    vout = m.picard_map(P, vin, v, w, S0)
    inc = vout - v
    mid = 0.5 * (v + vin)
    s = np.abs(w[:, None] * inc).sum()
    assert np.all(np.abs((w[:, None] * inc).sum(axis=0)) < 1e-14 * s)
    assert abs(np.sum(w[:, None] * mid * inc)) < 1e-14 * s * np.abs(mid).max()


def test_anderson_matches_picard_fixed_point():
    # At dt = 0.005 the implicit system has one root near the predictor; at
    # dt >= 0.01 damped Picard and Anderson can settle on different roots a
    # few 1e-3 apart (both with residual < tol), so the comparison uses 0.005.
    p = collop.config("gaussian_n4096", N_PARTICLES=600, DT=0.005)
    v, w = collop.initial_markers(p)
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(p.DT, m=5, nthreads=4)
    v1, _ = m.predictor(P, v, w, fc)
    va, sa, ra = m.step_anderson(P, v1, v, w, S0)
    Pp = O.default_params(p.DT, m=5, nthreads=4, use_anderson=0, damping=0.7)
    vp, sp, rp = m.step_anderson(Pp, v1, v, w, S0)
    assert sa.exit_reason == 0 and sp.exit_reason == 0
    assert np.max(np.abs(va - vp)) < 1e-9
    # the residual trace ends below the effective tolerance
    assert ra[-1] < max(1e-12 * np.linalg.norm(va), 1e-10) * 1.0001


def test_time_step_invariants(small):
    m, v, w, fc, S0 = small["m"], small["v"], small["w"], small["fc"], small["S0"]
    P = small["P"]
    (p1, p2), E0 = O.momentum_energy(v, w)
    vv, ff, S = v, fc, S0
    for _ in range(3):
        vv, ff, st = m.time_step(P, vv, w, ff, S)
        assert st.exit_reason == 0
        assert abs(st.energy - E0) < 1e-12 * E0
        assert abs(st.p1 - p1) < 1e-14 and abs(st.p2 - p2) < 1e-14
        assert st.entropy >= S - 1e-14
        S = st.entropy


def test_exits(small):
    m, v, w, S0 = small["m"], small["v"], small["w"], small["S0"]
    # iteration cap: returns the best iterate, reason 2
    P = O.default_params(small["p"].DT, m=5, nthreads=4, max_iter=3)
    out, st, res = m.step_anderson(P, small["v1"], v, w, S0)
    assert st.exit_reason == 2 and st.iters == 3
    assert st.residual == pytest.approx(min(res))
    # stagnation: an unreachable tolerance with a short window
    P = O.default_params(small["p"].DT, m=5, nthreads=4, tol=0.0, abs_floor=0.0, stag_window=5, stag_rel_tol=0.5)
    out, st, res = m.step_anderson(P, small["v1"], v, w, S0)
    assert st.exit_reason == 1 and st.iters % 5 == 0


def test_ring_window_permutation_invariance():
    """The Anderson least-squares step is invariant under a common column
    permutation of dF and dG (the ring cursor of solver.jl:386-394)."""
    rng = np.random.default_rng(9)
    n, h = 50, 4
    dF = rng.standard_normal((n, h))
    dG = rng.standard_normal((n, h))
    r = rng.standard_normal(n)

    def step(F, Gm):
        A = F.T @ F
        A += (1e-10 * np.mean(np.diag(A)) + 1e-30) * np.eye(h)
        return Gm @ np.linalg.solve(A, F.T @ r)

    perm = [2, 0, 3, 1]
    np.testing.assert_allclose(step(dF, dG), step(dF[:, perm], dG[:, perm]), rtol=1e-12)


def test_uniform_breakpoints_exact():
    for n in (16, 32, 64):
        bp = collop.config("bimodal_n65536", bp1=uniform_bp(-6, 6, n)).bp1
        assert np.all(np.diff(bp) == bp[1] - bp[0])


def test_defect_correction_finds_the_anderson_root():
    """step_defect! (solver.jl:569-634) restated: its fixed point is a fixed point
    of the full Picard map (Phi_k(v_k) = G(v_k)), so it must return a root of the
    same implicit system as step_anderson!, with fewer full maps."""
    v, w = collop.markers("gaussian", 512, 3)
    bp = uniform_bp(-6, 6, 12)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(0.005, m=5)
    v1, _ = m.predictor(P, v, w, fc)
    va, sa, _ = m.step_anderson(P, v1, v, w, S0)
    P.solver = 1
    vd, sd, res = m.step_defect(P, v1, v, w, S0)
    assert sa.exit_reason == 0 and sd.exit_reason == 0
    assert sd.iters < sa.iters and sd.inner > 0
    assert np.all(np.diff(res) < 0)                      # the outer residual falls every iteration
    r = np.linalg.norm(m.picard_map(P, vd, v, w, S0) - vd)
    assert r <= 20 * max(1e-12 * np.linalg.norm(vd), 1e-10)
    assert np.max(np.abs(vd - va)) <= 1e-10


def test_defect_extensions_keep_the_root():
    """Outer Anderson and the exact near field change the iteration, not the
    fixed point: the extended solve lands on the reference solve's root."""
    v, w = collop.markers("bimodal", 1024, 5, sigma1=0.5 ** 0.5, sigma2=0.5 ** 0.5, peak=2.0)
    bp = uniform_bp(-6, 6, 16)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(0.005, m=5, solver=1)
    v1, _ = m.predictor(P, v, w, fc)
    va, sa, _ = m.step_defect(P, v1, v, w, S0)
    P.dc_outer_m, P.dc_near_theta = 3, 0.5   # theta large enough to catch pairs at N = 1024
    vb, sb, _ = m.step_defect(P, v1, v, w, S0)
    assert sa.exit_reason == 0 and sb.exit_reason == 0
    assert np.max(np.abs(va - vb)) <= 1e-10
    r = np.linalg.norm(m.picard_map(P, vb, v, w, S0) - vb)
    assert r <= 20 * max(1e-12 * np.linalg.norm(vb), 1e-10)


def test_defect_handover_with_one_inner_map_is_the_anderson_solve():
    """The hand-over (dc_fallback, an extension) continues the step with
    step_anderson! from the outer iterate.  With one inner map per outer
    iteration no frozen-metric solve can reach its tolerance (at u = v_k the
    frozen map's residual is the outer one), so the step hands over after its
    first outer map and must return step_anderson!'s result from the same
    initial guess bit for bit, with one map more."""
    v, w = collop.markers("gaussian", 512, 3)
    bp = uniform_bp(-6, 6, 12)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(0.02, m=5)
    v1, _ = m.predictor(P, v, w, fc)
    va, sa, _ = m.step_anderson(P, v1, v, w, S0)
    Q = O.default_params(0.02, m=5, solver=1, dc_max_inner=1, dc_fallback=1)
    vd, sd, _ = m.step_defect(Q, v1, v, w, S0)
    assert sd.handover == 1 and sa.handover == 0
    assert sd.iters == sa.iters + 1 and sd.exit_reason == sa.exit_reason == 0
    np.testing.assert_array_equal(vd, va)
    Q.dc_fallback = 0                                 # without the hand-over: the reference's outer iteration
    ve, se, _ = m.step_defect(Q, v1, v, w, S0)
    assert se.handover == 0 and se.inner == se.iters - (se.exit_reason == 0)


def test_defect_handover_lands_on_a_root():
    """A step whose first frozen-metric solve needs more than dc_max_inner maps
    (isotropic N = 1024, dt = 0.05) hands over and converges: the result is a
    root of the full Picard map within the tolerance, with momentum exact and
    energy to the tolerance's order."""
    v, w = collop.markers("gaussian", 1024, 42)
    bp = uniform_bp(-6, 6, 16)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(0.05, m=5, solver=1, dc_outer_m=3, dc_near_theta=0.02, dc_max_inner=10, dc_fallback=1)
    vn, fcn, st = m.time_step(P, v, w, fc, S0)
    assert st.handover >= 1 and st.exit_reason == 0
    r = np.linalg.norm(m.picard_map(P, vn, v, w, S0) - vn)
    assert r <= 20 * max(1e-12 * np.linalg.norm(vn), 1e-10)
    (p1, p2), E0 = O.momentum_energy(v, w)
    (q1, q2), E1 = O.momentum_energy(vn, w)
    assert abs(q1 - p1) <= 1e-14 and abs(q2 - p2) <= 1e-14
    assert abs(E1 - E0) <= 1e-11 * E0
    assert m.entropy(fcn) >= S0


def test_defect_handover_needs_an_iteration_left():
    """With max_iter = 1 there is no iteration left for the Anderson solve: the
    step leaves through the iteration cap after its one outer map, without a
    hand-over, as step_defect! does."""
    v, w = collop.markers("gaussian", 512, 3)
    bp = uniform_bp(-6, 6, 12)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(0.02, m=5, solver=1, dc_max_inner=1, dc_fallback=1, max_iter=1)
    v1, _ = m.predictor(P, v, w, fc)
    vd, sd, _ = m.step_defect(P, v1, v, w, S0)
    assert sd.handover == 0 and sd.exit_reason == 2 and sd.iters == 1
    P.max_iter = 2                                   # one iteration left: the hand-over runs one Anderson map
    vd, sd, _ = m.step_defect(P, v1, v, w, S0)
    assert sd.handover == 1 and sd.iters == 2 and sd.exit_reason == 2
