"""Pins the oracle's matrix-free GMRES and Jacobian-free Newton-Krylov
restatement (newton_krylov.jl) with the cases of the reference's own solver
tests (test/integration/newton_krylov.jl:10-71), checked against numpy, and
pins step_newton! (solver.jl:437-486) by the root the Anderson step finds.
"""
import numpy as np

import collop
import oracle as O


def test_gmres_matches_direct_solve():
    rng = np.random.default_rng(7)
    n = 60
    A = np.eye(n) + 0.3 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, k, r = O.gmres_dense(A, b, rtol=1e-12)
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    assert abs(r - np.linalg.norm(A @ x - b)) <= 1e-10 * np.linalg.norm(b)
    assert k <= n
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-10)
    # a loose tolerance stops early with an honest residual
    x2, k2, r2 = O.gmres_dense(A, b, rtol=1e-2)
    assert k2 < k and r2 <= 1e-2 * np.linalg.norm(b)
    assert abs(np.linalg.norm(A @ x2 - b) - r2) <= 1e-6 * r2
    # the iteration cap is respected
    assert O.gmres_dense(A, b, rtol=1e-14, itmax=3)[1] == 3
    # zero right-hand side: no products, zero solution
    x0, k0, r0 = O.gmres_dense(A, np.zeros(n), rtol=1e-8)
    assert (k0, r0) == (0, 0.0) and not x0.any()


def test_gmres_minimises_over_the_krylov_space():
    # after k products the iterate minimises ||A x - b|| over span{b, ..., A^{k-1} b}
    rng = np.random.default_rng(9)
    n, k = 30, 6
    A = np.eye(n) + 0.5 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x, kk, r = O.gmres_dense(A, b, rtol=1e-16, itmax=k)
    assert kk == k
    K = np.empty((n, k))
    K[:, 0] = b
    for j in range(1, k):
        K[:, j] = A @ K[:, j - 1]
    Q, _ = np.linalg.qr(K)
    y, *_ = np.linalg.lstsq(A @ Q, b, rcond=None)
    np.testing.assert_allclose(x, Q @ y, rtol=0, atol=1e-10)
    assert abs(r - np.linalg.norm(A @ x - b)) < 1e-10


def _midpoint(seed=11, n=40, dt=0.2):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n)) / np.sqrt(n)
    x0 = rng.standard_normal(n)
    f = lambda y: np.tanh(B @ y) - 0.5 * y ** 3            # noqa: E731
    F = lambda x: x - x0 - dt * f(0.5 * (x0 + x))           # noqa: E731
    return B, x0, dt, f, F


def test_newton_krylov_midpoint_fixed_point():
    B, x0, dt, f, F = _midpoint()
    xg = x0 + dt * f(x0)                  # explicit Euler predictor
    x, Fx, st, ne, rn, nn, r0 = O.newton_midpoint(B, x0, dt, xg, tol=1e-11, max_evals=200, fd_h=1e-7)
    assert st == 0 and rn < 1e-11
    assert abs(np.linalg.norm(F(x)) - rn) < 1e-12
    np.testing.assert_allclose(Fx, F(x), rtol=0, atol=1e-12)
    assert r0 > rn and nn <= 6 and ne < 60
    # budget exhaustion returns the best iterate and says so
    x, Fx, st, ne, rb, _, _ = O.newton_midpoint(B, x0, dt, xg, tol=1e-14, max_evals=4, fd_h=1e-7)
    assert st == 2 and ne <= 4
    assert abs(np.linalg.norm(F(x)) - rb) < 1e-12 * max(1.0, rb)


def test_step_newton_finds_the_anderson_root():
    p = collop.config("gaussian_n4096", N_PARTICLES=600, DT=0.005)
    v, w = collop.initial_markers(p)
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    Pa = O.default_params(p.DT, m=5, nthreads=4)
    Pn = O.default_params(p.DT, m=5, nthreads=4, solver=2)
    va, _, sa = m.time_step(Pa, v, w, fc, S0)
    vn, _, sn = m.time_step(Pn, v, w, fc, S0)
    assert sa.exit_reason == 0 and sn.exit_reason == 0
    assert np.max(np.abs(va - vn)) < 1e-9
    (p1, p2), E0 = O.momentum_energy(v, w)
    assert abs(sn.p1 - p1) < 1e-14 and abs(sn.p2 - p2) < 1e-14
    assert abs(sn.energy - E0) < 1e-12 * E0 and sn.entropy >= S0
    assert sn.n_restart >= 1          # Newton steps taken
