"""Pins the oracle's Lenard-Bernstein restatement (functions.jl:66-115,
solver.jl:185-222, main.jl:233-238) against mathematics it must satisfy:
the 3x3 multipliers against numpy, exact momentum and energy conservation of
the right-hand side, the Maxwellian equilibrium in closed form, the Gonzalez
discrete chain rule of the LB G_eff, and the invariants of an LB step.
"""
import numpy as np
import pytest

import collop
import oracle as O


def _rand(N, seed):
    rng = np.random.default_rng(seed)
    v = np.asfortranarray(rng.standard_normal((N, 2)) * [1.3, 0.7] + [0.4, -0.2])
    w = rng.uniform(0.5, 1.5, N) / N
    g = np.asfortranarray(rng.standard_normal((N, 2)))
    return v, w, g


def test_lb_multipliers_vs_numpy_and_conservation():
    v, w, g = _rand(500, 1)
    nu = 1.7
    dot, AB = O.lb_rhs(v, w, g, nu)
    n, U, Q = w.sum(), w @ v, w @ (v ** 2).sum(axis=1)
    M = np.array([[n, 0, U[0]], [0, n, U[1]], [U[0], U[1], Q]])
    rhs = -np.r_[w @ g, w @ (v * g).sum(axis=1)]
    np.testing.assert_allclose(AB, np.linalg.solve(M, rhs), rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(dot, -nu * (g + AB[:2] + AB[2] * v), rtol=0, atol=1e-14)
    # momentum and energy of the drift vanish: sum w v_dot = 0, sum w v.v_dot = 0
    s = np.abs(w[:, None] * dot).sum()
    assert np.all(np.abs((w[:, None] * dot).sum(axis=0)) < 1e-14 * s)
    assert abs(np.sum(w[:, None] * v * dot)) < 1e-14 * s * np.abs(v).max()


def test_lb_maxwellian_equilibrium():
    v, w, _ = _rand(400, 2)
    u, T = np.array([0.3, -0.1]), 1.7
    g = np.asfortranarray(-(v - u) / T)       # grad log of a Maxwellian (u, T)
    dot, AB = O.lb_rhs(v, w, g, 2.0)
    np.testing.assert_allclose(AB, np.r_[-u / T, 1.0 / T], rtol=1e-12, atol=1e-14)
    assert np.abs(dot).max() < 1e-13


@pytest.fixture(scope="module")
def lb_small():
    p = collop.config("lb_bimodal_n40000", N_PARTICLES=800, DT=0.001)
    v, w = collop.initial_markers(p)
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(p.DT, m=8, nthreads=4, collision_model=1, nu=1.0)
    v1, _ = m.predictor(P, v, w, fc)
    return dict(p=p, v=v, w=w, m=m, fc=fc, S0=S0, P=P, v1=v1)


def test_lb_predictor_is_explicit_euler(lb_small):
    m, v, w, fc, P = lb_small["m"], lb_small["v"], lb_small["w"], lb_small["fc"], lb_small["P"]
    g = m.loggrad(v, fc, fs_floor=1e-30, g_max=100.0)
    dot, _ = O.lb_rhs(v, w, g, 1.0)
    np.testing.assert_allclose(lb_small["v1"], v + P.dt * dot, rtol=0, atol=1e-15)


def test_lb_gonzalez_discrete_chain_rule(lb_small):
    m, v, w, S0 = lb_small["m"], lb_small["v"], lb_small["w"], lb_small["S0"]
    vin = lb_small["v1"]
    Geff = m.lb_geff(vin, v, w, S0)
    S1 = m.entropy(m.l2_project(vin, w))
    lhs = np.sum(-w[:, None] * Geff * (vin - v))
    assert abs(lhs - (S1 - S0)) < 1e-13 * max(1.0, abs(S1 - S0)) + 1e-15
    g0 = m.lb_geff(vin, v, w, S0, use_gonzalez=False)
    g_direct = m.loggrad(0.5 * (v + vin), m.l2_project(0.5 * (v + vin), w), fs_floor=1e-30, g_max=100.0)
    np.testing.assert_array_equal(g0, g_direct)


def test_lb_step_invariants(lb_small):
    m, v, w, fc, S0, P = (lb_small[k] for k in ("m", "v", "w", "fc", "S0", "P"))
    (p1, p2), E0 = O.momentum_energy(v, w)
    vv, ff, S = v, fc, S0
    for _ in range(3):
        vv, ff, st = m.time_step(P, vv, w, ff, S)
        assert st.exit_reason == 0
        assert abs(st.p1 - p1) < 1e-14 and abs(st.p2 - p2) < 1e-14
        assert abs(st.energy - E0) < 1e-12 * E0
        assert st.entropy >= S - 1e-14
        S = st.entropy
    # the state relaxes: the LB drift moves markers (unlike a zero operator)
    assert np.abs(vv - v).max() > 1e-3
