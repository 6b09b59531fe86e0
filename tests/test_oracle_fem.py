"""Pins the oracle's FEM half (MantisWrappers.jl:90-289, functions.jl:37-64,
121-202) against independent mathematics: numpy Gauss-Legendre, scipy
B-splines, closed-form extraction, Kronecker structure, moment reproduction,
exact gradients of polynomial fields and analytic integrals.
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle as O
from conftest import uniform_bp

NONUNI = np.array([-6.0, -5.0, -4.0, -3.0, -1.5, 0.0, 0.7, 2.0, 4.0, 5.5, 6.0])


def knots(bp, p, k):
    mult = p - k
    return np.r_[[bp[0]] * (p + 1), np.repeat(bp[1:-1], mult), [bp[-1]] * (p + 1)]


def bernstein(p, xi):
    from math import comb
    return np.array([comb(p, i) * xi ** i * (1 - xi) ** (p - i) for i in range(p + 1)])


@pytest.mark.parametrize("n", range(1, 11))
def test_gauss_legendre(n):
    x, w = O.gauss_legendre01(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, (xr + 1) / 2, atol=2e-16 * 4)
    np.testing.assert_allclose(w, wr / 2, atol=1e-15)


@pytest.mark.parametrize("bp,p,k", [(uniform_bp(-6, 6, 16), 2, 1), (NONUNI, 2, 1), (NONUNI, 3, 2),
                                    (NONUNI, 2, 0), (uniform_bp(-1, 1, 1), 2, 1), (uniform_bp(0, 3, 3), 3, 1)])
def test_extraction_vs_scipy(bp, p, k):
    C, bs, nd = O.extraction_1d(bp, p, k)
    t = knots(bp, p, k)
    assert nd == len(t) - p - 1
    for e in range(len(bp) - 1):
        for xi in np.linspace(0.03, 0.97, 9):
            x = bp[e] + xi * (bp[e + 1] - bp[e])
            phi = C[e].T @ bernstein(p, xi)
            for j in range(p + 1):
                c = np.zeros(nd)
                c[bs[e] + j] = 1.0
                assert abs(BSpline(t, c, p)(x) - phi[j]) < 1e-14
            # every basis function not listed for this cell vanishes there
            allv = np.array([BSpline(t, np.eye(nd)[i], p)(x) for i in range(nd)])
            assert abs(allv.sum() - 1.0) < 1e-14
            assert np.all(np.abs(np.delete(allv, range(bs[e], bs[e] + p + 1))) < 1e-15)


def test_extraction_closed_form_uniform_p2():
    C, bs, nd = O.extraction_1d(uniform_bp(-6, 6, 8), 2, 1)
    first = np.array([[1, 0, 0], [0, 1, 0], [0, .5, .5]])
    inner = np.array([[.5, .5, 0], [0, 1, 0], [0, .5, .5]])
    last = np.array([[.5, .5, 0], [0, 1, 0], [0, 0, 1]])
    np.testing.assert_array_equal(C[0], first)
    for e in range(1, 7):
        np.testing.assert_array_equal(C[e], inner)
    np.testing.assert_array_equal(C[7], last)
    np.testing.assert_array_equal(bs, np.arange(8))
    assert nd == 10


def mass_1d_scipy(bp, p=2, k=1):
    t = knots(bp, p, k)
    nd = len(t) - p - 1
    xg, wg = np.polynomial.legendre.leggauss(8)
    M = np.zeros((nd, nd))
    for e in range(len(bp) - 1):
        a, b = bp[e], bp[e + 1]
        x = a + (xg + 1) / 2 * (b - a)
        B = np.array([BSpline(t, np.eye(nd)[i], p)(x) for i in range(nd)])
        M += (B * (wg / 2 * (b - a))) @ B.T
    return M


@pytest.mark.parametrize("bp1,bp2", [(uniform_bp(-6, 6, 8), uniform_bp(-6, 6, 8)), (NONUNI, uniform_bp(-3, 5, 5))])
def test_mass_matrix_kronecker(bp1, bp2):
    m = O.Mesh(bp1, bp2)
    M = m.mass_dense()
    M1, M2 = mass_1d_scipy(bp1), mass_1d_scipy(bp2)
    # dof index i1 + nd1 i2 -> M = M2 (x) M1 in numpy's kron ordering
    np.testing.assert_allclose(M, np.kron(M2, M1), atol=1e-14 * np.abs(M).max())
    b = np.random.default_rng(1).standard_normal(m.n_dofs)
    x = m.mass_solve(b)
    np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=0, atol=1e-12 * np.abs(x).max())
    assert np.max(np.abs(M @ x - b)) < 1e-13 * np.abs(b).max() * np.linalg.cond(M)


def test_l2_rhs_partition_of_unity_and_domain():
    bp = uniform_bp(-3, 3, 6)
    m = O.Mesh(bp, bp)
    rng = np.random.default_rng(2)
    v = np.asfortranarray(rng.uniform(-4, 4, (500, 2)))
    w = rng.uniform(0.5, 1.5, 500)
    rhs = m.l2_rhs(v, w)
    inside = (np.abs(v[:, 0]) < 3) & (np.abs(v[:, 1]) < 3)
    assert abs(rhs.sum() - w[inside].sum()) < 1e-12 * w.sum()
    assert np.all(m.l2_rhs(v[~inside], w[~inside]) == 0.0)
    assert np.all(m.l2_rhs(np.zeros((0, 2)), np.zeros(0)) == 0.0)


def test_projection_reproduces_moments():
    # Quadratics lie in the space, so the L2 projection keeps sum w p(v) for p in
    # {1, v1, v2, v1^2, v2^2, v1 v2}: int f_s p = f^T M c_p with c_p the spline
    # coefficients of p (by blossoming: x -> (t_{k+1}+t_{k+2})/2, x^2 -> t_{k+1} t_{k+2}).
    bp1, bp2 = NONUNI, uniform_bp(-4, 4, 7)
    m = O.Mesh(bp1, bp2)
    v = O.colmajor(np.random.default_rng(3).normal(0, 1.2, (3000, 2)))
    w = np.full(3000, 1 / 3000)
    fc = m.l2_project(v, w)
    M = m.mass_dense()
    t1, t2 = knots(bp1, 2, 1), knots(bp2, 2, 1)
    g1 = (t1[1:-2] + t1[2:-1]) / 2
    g2 = (t2[1:-2] + t2[2:-1]) / 2
    s1 = t1[1:-2] * t1[2:-1]
    s2 = t2[1:-2] * t2[2:-1]
    one1, one2 = np.ones_like(g1), np.ones_like(g2)
    inside = (v[:, 0] > bp1[0]) & (v[:, 0] < bp1[-1]) & (v[:, 1] > bp2[0]) & (v[:, 1] < bp2[-1])
    vi, wi = v[inside], w[inside]
    for c1, c2, p in [(one1, one2, np.ones(len(vi))), (g1, one2, vi[:, 0]), (one1, g2, vi[:, 1]),
                      (s1, one2, vi[:, 0] ** 2), (one1, s2, vi[:, 1] ** 2), (g1, g2, vi[:, 0] * vi[:, 1])]:
        cp = np.kron(c2, c1)
        assert abs(fc @ M @ cp - np.sum(wi * p)) < 1e-13


def _coeffs_of(bp1, bp2, f1, f2):
    """Spline coefficients of f1(x) f2(y) for f in {1, x, x^2} via blossoms."""
    def c(bp, f):
        t = knots(bp, 2, 1)
        a, b = t[1:-2], t[2:-1]
        return {"1": np.ones_like(a), "x": (a + b) / 2, "xx": a * b}[f]
    return np.kron(c(bp2, f2), c(bp1, f1))


def test_compute_G_exact_for_polynomials():
    bp1, bp2 = NONUNI, uniform_bp(-6, 6, 9)
    m = O.Mesh(bp1, bp2)
    v = O.colmajor(np.random.default_rng(4).uniform(-7, 7, (400, 2)))
    inside = (np.abs(v[:, 0]) < 6) & (np.abs(v[:, 1]) < 6)
    L = 2.0 * _coeffs_of(bp1, bp2, "x", "1") - 3.0 * _coeffs_of(bp1, bp2, "1", "x") + 0.5 * _coeffs_of(bp1, bp2, "1", "1")
    G = m.compute_G(v, L)
    np.testing.assert_allclose(G[inside], np.tile([2.0, -3.0], (inside.sum(), 1)), atol=1e-12)
    assert np.all(G[~inside] == 0.0)
    L = _coeffs_of(bp1, bp2, "xx", "1") + _coeffs_of(bp1, bp2, "x", "x")
    G = m.compute_G(v, L)
    exp = np.c_[2 * v[:, 0] + v[:, 1], v[:, 0]]
    np.testing.assert_allclose(G[inside], exp[inside], atol=1e-11)


def test_loggrad_exact_and_clamped():
    bp = uniform_bp(-6, 6, 12)
    m = O.Mesh(bp, bp)
    fc = _coeffs_of(bp, bp, "1", "1") + 0.1 * _coeffs_of(bp, bp, "xx", "1")   # f = 1 + 0.1 v1^2
    v = O.colmajor(np.random.default_rng(5).uniform(-6.5, 6.5, (300, 2)))
    inside = (np.abs(v[:, 0]) < 6) & (np.abs(v[:, 1]) < 6)
    g = m.loggrad(v, fc)
    exp = np.c_[0.2 * v[:, 0] / (1 + 0.1 * v[:, 0] ** 2), np.zeros(300)]
    np.testing.assert_allclose(g[inside], exp[inside], atol=1e-13)
    assert np.all(g[~inside] == 0.0)
    # a tiny f makes grad f / f explode: clamped to +-g_max (functions.jl:59-61)
    g = m.loggrad(v, 1e-20 * fc + 1e-3 * _coeffs_of(bp, bp, "x", "1"), fs_floor=1e-30, g_max=100.0)
    assert np.all(np.abs(g) <= 100.0)


def test_entropy_and_r_constant_field():
    bp1, bp2 = NONUNI, uniform_bp(-2, 2, 4)
    m = O.Mesh(bp1, bp2)
    c0 = 0.37
    fc = np.full(m.n_dofs, c0)
    area = (bp1[-1] - bp1[0]) * (bp2[-1] - bp2[0])
    assert abs(m.entropy(fc) - (-c0 * np.log(c0) * area)) < 1e-13
    r = m.compute_r(fc)
    intphi = m.mass_dense().sum(axis=1)   # sum_j phi_j = 1  =>  int phi_i = row sum of M
    np.testing.assert_allclose(r, (1 + np.log(c0)) * intphi, atol=1e-14)
    # below the floor nothing contributes; the logsq variant keeps |f|
    assert m.entropy(np.full(m.n_dofs, -c0)) == 0.0
    assert abs(m.entropy(np.full(m.n_dofs, -c0), use_logsq=True) - (c0 * np.log(c0) * area)) < 1e-13
    assert np.all(m.compute_r(np.full(m.n_dofs, -c0)) == 0.0)


def test_entropy_and_r_linear_field_vs_fine_quadrature():
    bp = uniform_bp(-2, 2, 5)
    m = O.Mesh(bp, bp)
    fc = 1.5 * _coeffs_of(bp, bp, "1", "1") + 0.3 * _coeffs_of(bp, bp, "x", "1") + 0.2 * _coeffs_of(bp, bp, "1", "x")
    f = lambda x, y: 1.5 + 0.3 * x + 0.2 * y          # noqa: E731  (>= 0.5 on [-2,2]^2)
    # S = -int f log f, per cell with a 24-point tensor Gauss rule
    xg, wg = np.polynomial.legendre.leggauss(24)
    t = knots(bp, 2, 1)
    nd = len(t) - 3
    S = 0.0
    r = np.zeros((nd, nd))
    for i in range(len(bp) - 1):
        for j in range(len(bp) - 1):
            x = bp[i] + (xg + 1) / 2 * (bp[i + 1] - bp[i])
            y = bp[j] + (xg + 1) / 2 * (bp[j + 1] - bp[j])
            W = np.outer(wg, wg) / 4 * (bp[i + 1] - bp[i]) * (bp[j + 1] - bp[j])
            F = f(x[:, None], y[None, :])
            S -= np.sum(W * F * np.log(F))
            Bx = np.array([BSpline(t, np.eye(nd)[k], 2)(x) for k in range(nd)])
            By = np.array([BSpline(t, np.eye(nd)[k], 2)(y) for k in range(nd)])
            r += Bx @ (W * (1 + np.log(F))) @ By.T
    assert abs(m.entropy(fc) - S) < 1e-10
    np.testing.assert_allclose(m.compute_r(fc), r.T.reshape(-1), atol=1e-10)


def test_eval_particle_matches_scipy():
    bp1, bp2 = NONUNI, uniform_bp(-6, 6, 9)
    m = O.Mesh(bp1, bp2)
    t1, t2 = knots(bp1, 2, 1), knots(bp2, 2, 1)
    for (x, y) in [(0.31, -2.2), (-5.9, 5.99), (3.3, 0.0), (5.5, 1.0)]:
        vals, dx1, dx2, gids, h = m.eval_particle(x, y)
        for k in range(9):
            i1, i2 = gids[k] % m.nd1, gids[k] // m.nd1
            b1 = BSpline(t1, np.eye(m.nd1)[i1], 2)
            b2 = BSpline(t2, np.eye(m.nd2)[i2], 2)
            assert abs(vals[k] - b1(x) * b2(y)) < 1e-14
            assert abs(dx1[k] / h[0] - b1.derivative()(x) * b2(y)) < 1e-13
            assert abs(dx2[k] / h[1] - b1(x) * b2.derivative()(y)) < 1e-13
    assert m.eval_particle(6.0, 0.0) is None and m.eval_particle(0.0, -6.0) is None


# ---- run-record diagnostics (functions.jl:300-343) ---------------------------

def _gauss_sum(bp1, bp2, fc, fn):
    """sum over cells and 6x6 Gauss points of fn(f, cell) w_q |e|, f from scipy
    B-splines: an implementation independent of the oracle's extraction."""
    xg, wg = np.polynomial.legendre.leggauss(6)
    t1, t2 = knots(bp1, 2, 1), knots(bp2, 2, 1)
    nd1, nd2 = len(t1) - 3, len(t2) - 3
    C = np.asarray(fc).reshape(nd2, nd1)           # gid = i1 + nd1 * i2
    tot = 0.0
    for j in range(len(bp2) - 1):
        for i in range(len(bp1) - 1):
            x = bp1[i] + (xg + 1) / 2 * (bp1[i + 1] - bp1[i])
            y = bp2[j] + (xg + 1) / 2 * (bp2[j + 1] - bp2[j])
            jac = (bp1[i + 1] - bp1[i]) * (bp2[j + 1] - bp2[j])
            Bx = np.array([BSpline(t1, np.eye(nd1)[k], 2)(x) for k in range(nd1)])
            By = np.array([BSpline(t2, np.eye(nd2)[k], 2)(y) for k in range(nd2)])
            F = By.T @ C @ Bx                        # F[b, a] = f(x_a, y_b)
            W = np.outer(wg, wg).T / 4               # W[b, a] = w_a w_b
            tot += np.sum(fn(F, i, j) * W) * jac
    return tot


def test_negative_part_closed_forms():
    bp1, bp2 = NONUNI, uniform_bp(-2, 2, 4)
    m = O.Mesh(bp1, bp2)
    area = (bp1[-1] - bp1[0]) * (bp2[-1] - bp2[0])
    assert m.negative_part_l1(np.full(m.n_dofs, 0.4)) == 0.0
    assert abs(m.negative_part_l1(np.full(m.n_dofs, -0.4)) - 0.4 * area) < 1e-13
    # f = x on [-2, 2]^2 with a breakpoint at the kink: int max(-x, 0) = 2 * 4
    bp = uniform_bp(-2, 2, 4)
    m2 = O.Mesh(bp, bp)
    assert abs(m2.negative_part_l1(_coeffs_of(bp, bp, "x", "1")) - 8.0) < 1e-13


def test_negative_part_vs_scipy_gauss_sum():
    rng = np.random.default_rng(5)
    bp1, bp2 = NONUNI, uniform_bp(-3, 3, 7)
    m = O.Mesh(bp1, bp2)
    fc = rng.standard_normal(m.n_dofs) * 0.3 + 0.1
    ref = _gauss_sum(bp1, bp2, fc, lambda F, i, j: np.maximum(-F, 0.0))
    assert ref > 0.1
    assert abs(m.negative_part_l1(fc) - ref) < 1e-13 * max(1.0, ref)


def test_fs_minus_fp_closed_forms():
    bp1, bp2 = NONUNI, uniform_bp(-2, 2, 4)
    m = O.Mesh(bp1, bp2)
    c = 0.25
    fc = np.full(m.n_dofs, c)                    # f_s = c (partition of unity)
    h1, h2 = np.diff(bp1), np.diff(bp2)
    cx = (bp1[:-1] + bp1[1:]) / 2
    cy = (bp2[:-1] + bp2[1:]) / 2
    X, Y = np.meshgrid(cx, cy, indexing="xy")     # one marker per cell, mass c |e|
    v = np.c_[X.ravel(), Y.ravel()]
    w = (c * np.outer(h2, h1)).ravel()
    assert m.fs_minus_fp_l2(fc, v, w) < 1e-14
    # markers outside the open box (and on its boundary) are not counted
    v_out = np.r_[v, [[bp1[0], 0.0], [9.0, 0.1], [0.2, bp2[-1]]]]
    w_out = np.r_[w, [5.0, 7.0, 3.0]]
    assert m.fs_minus_fp_l2(fc, v_out, w_out) < 1e-14
    # move marker 0's mass into cell 1: |d| = dm / |e| on both cells
    dm = 0.3 * w[0]
    w2 = w.copy()
    w2[0] -= dm
    w2[1] += dm
    e0, e1 = h1[0] * h2[0], h1[1] * h2[0]
    exact = np.sqrt(dm ** 2 / e0 + dm ** 2 / e1)
    assert abs(m.fs_minus_fp_l2(fc, v, w2) - exact) < 1e-14 * max(1.0, exact) + 1e-15


def test_fs_minus_fp_vs_numpy_histogram():
    rng = np.random.default_rng(8)
    bp1, bp2 = NONUNI, uniform_bp(-3, 3, 7)
    m = O.Mesh(bp1, bp2)
    N = 3000
    v = np.c_[rng.uniform(-7, 7, N), rng.uniform(-3.5, 3.5, N)]   # some outside
    w = rng.uniform(0.5, 1.5, N) / N
    fc = m.l2_project(v, w)
    inside = (v[:, 0] > bp1[0]) & (v[:, 0] < bp1[-1]) & (v[:, 1] > bp2[0]) & (v[:, 1] < bp2[-1])
    i = np.searchsorted(bp1, v[inside, 0], side="right") - 1
    j = np.searchsorted(bp2, v[inside, 1], side="right") - 1
    ne1 = len(bp1) - 1
    mass = np.bincount(i + ne1 * j, weights=w[inside], minlength=ne1 * (len(bp2) - 1))
    jac = np.outer(np.diff(bp2), np.diff(bp1)).ravel()
    fp = mass / jac
    ref = np.sqrt(_gauss_sum(bp1, bp2, fc, lambda F, ii, jj: (F - fp[ii + ne1 * jj]) ** 2))
    assert ref > 1e-3
    assert abs(m.fs_minus_fp_l2(fc, v, w) - ref) < 1e-12 * ref
