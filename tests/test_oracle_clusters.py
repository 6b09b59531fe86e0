"""The oracle's stiff-cluster elimination (an extension of step_anderson!, no
counterpart in the reference) against what fixes it independently:

  * grad L and Hess L of the spline against differences of compute_G;
  * the analytic Jacobian of a cluster's midpoint equations against central
    differences of the equations;
  * the cluster finder against a brute-force pair search, and its size cap;
  * a pair's root lies on the circle over a = v0_g - v0_a + dt/2 (ext_g - ext_a)
    (d . (d - a) = 0) and solves the pair's equations; larger clusters' roots
    solve theirs;
  * on a stiff state (close pairs formed, tests/golden/stiff_bimodal_n8192.npz),
    the step converges where the plain Anderson solve stagnates, and the result
    is a root of the plain Picard map (the reference's fixed point)."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import uniform_bp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stiff_bimodal_n8192.npz")


@pytest.fixture(scope="module")
def mesh():
    bp = uniform_bp(-6, 6, 32)
    return O.Mesh(bp, bp)


def field(mesh, N=4096, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((N, 2)) * 0.7
    w = np.full(N, 1.0 / N)
    fc = mesh.l2_project(v, w)
    return mesh.mass_solve(mesh.compute_r(fc))


def test_grad_hess_against_compute_G(mesh):
    L = field(mesh)
    h = 1e-6
    for x, y in ((0.31, -0.77), (-1.9, 0.05), (2.2, 1.3)):
        g, H = mesh.grad_hess(L, x, y)
        G = mesh.compute_G(np.array([[x, y]]), L)[0]
        np.testing.assert_allclose(g, G, rtol=0, atol=1e-13)
        Gx = (mesh.compute_G(np.array([[x + h, y]]), L)[0] - mesh.compute_G(np.array([[x - h, y]]), L)[0]) / (2 * h)
        Gy = (mesh.compute_G(np.array([[x, y + h]]), L)[0] - mesh.compute_G(np.array([[x, y - h]]), L)[0]) / (2 * h)
        np.testing.assert_allclose([H[0], H[1]], Gx, rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose([H[1], H[2]], Gy, rtol=1e-6, atol=1e-6)
    assert mesh.grad_hess(L, 6.5, 0.0) is None


def test_cluster_jacobian_against_differences(mesh):
    L = field(mesh)
    rng = np.random.default_rng(1)
    k = 4
    v0 = np.array([[0.3, -0.7]]) + rng.standard_normal((k, 2)) * 1e-5
    w = np.full(k, 1.0 / 4096)
    ext = rng.standard_normal((k, 2)) * 0.1
    m = v0 + rng.standard_normal((k, 2)) * 3e-6
    for t in (0.3, 1.0):
        F, J = mesh.cluster_residual(L, -7e-5, 0.02, v0, w, ext, m, t, jacobian=True)
        Jfd = np.zeros_like(J)
        for c in range(2 * k):
            e = np.zeros(2 * k)
            e[c] = 1e-10
            Jfd[:, c] = (mesh.cluster_residual(L, -7e-5, 0.02, v0, w, ext, m.ravel() + e, t)
                         - mesh.cluster_residual(L, -7e-5, 0.02, v0, w, ext, m.ravel() - e, t)) / 2e-10
        assert np.max(np.abs(J - Jfd)) <= 1e-5 * np.max(np.abs(J))


def brute_links(v, w, box, theta):
    out = set()
    N = v.shape[0]
    inb = (v[:, 0] > box[0]) & (v[:, 0] < box[1]) & (v[:, 1] > box[2]) & (v[:, 1] < box[3])
    for i in range(N):
        for j in range(i + 1, N):
            q = float(np.sum((v[i] - v[j]) ** 2))
            if inb[i] and inb[j] and q >= 1e-24 and theta * theta * q < max(w[i], w[j]) ** 2:
                out.add((i, j))
    return out


def test_cluster_finder_and_cap(mesh):
    rng = np.random.default_rng(3)
    N = 600
    v = rng.uniform(-5, 5, (N, 2))
    w = np.full(N, 1.0 / N)
    theta = 0.5                      # links within 1 / (N theta) = 3.3e-3
    # planted: a pair, a triple and a chain of five, spaced 1e-3; one pair outside the box
    v[1] = v[0] + [1e-3, 0.0]
    v[11] = v[10] + [0.0, 1e-3]
    v[12] = v[10] + [1e-3, 1e-3]
    for s in range(1, 5):
        v[20 + s] = v[20] + [2e-3 * s, 0.0]
    v[30] = [5.9995, 0.0]
    v[31] = [6.0005, 0.0]
    links = brute_links(v, w, mesh.box, theta)
    cl = mesh.clusters(v, w, theta, 32)
    got = {tuple(sorted((int(a), int(b)))) for c in cl for a in c for b in c if a < b}
    # every brute-force link joins two members of one cluster
    assert links <= got
    members = sorted(int(a) for c in cl for a in c)
    assert members == sorted({a for l in links for a in l})
    assert [0, 1] in [list(c) for c in cl] and [10, 11, 12] in [list(c) for c in cl]
    assert 30 not in members and 31 not in members
    # with a cap of 3 the chain of five splits into pieces of at most 3
    cl3 = mesh.clusters(v, w, theta, 3)
    assert max(len(c) for c in cl3) == 3
    assert sorted(int(a) for c in cl3 for a in c if 20 <= a <= 24) == list(range(20, 25))


def test_pair_and_cluster_roots(mesh):
    L = field(mesh)
    dt, lam = 0.02, -7e-5
    w2 = np.full(2, 1.0 / 4096)
    v0 = np.array([[0.4, -0.6], [0.4 + 2e-6, -0.6 + 1e-6]])
    ext = np.array([[0.03, -0.01], [0.028, -0.012]])
    m, _ = mesh.cluster_root(L, lam, dt, v0, w2, ext)
    assert m is not None
    F = mesh.cluster_residual(L, lam, dt, v0, w2, ext, m)
    assert np.max(np.abs(F)) <= 1e-13
    a = v0[0] - v0[1] + 0.5 * dt * (ext[0] - ext[1])
    d = m[0] - m[1]
    assert abs(np.dot(d, d - a)) <= 1e-12 * np.dot(a, a)
    # momentum: the weighted centre moves with ext alone
    np.testing.assert_allclose((w2[:, None] * m).sum(0), (w2[:, None] * (v0 + 0.5 * dt * ext)).sum(0), atol=1e-15)
    rng = np.random.default_rng(5)
    k = 5
    v0 = np.array([[-0.8, 0.9]]) + rng.standard_normal((k, 2)) * 2e-5
    wk = np.full(k, 1.0 / 4096)
    ext = 0.03 + rng.standard_normal((k, 2)) * 1e-3
    m, its = mesh.cluster_root(L, lam, dt, v0, wk, ext)
    assert m is not None and its > 0
    assert np.max(np.abs(mesh.cluster_residual(L, lam, dt, v0, wk, ext, m))) <= 1e-12 * (np.abs(v0).max() + 1)
    np.testing.assert_allclose((wk[:, None] * m).sum(0), (wk[:, None] * (v0 + 0.5 * dt * ext)).sum(0), atol=1e-15)


def test_stiff_step_converges_on_a_root(mesh):
    z = np.load(GOLD)
    v, w = O.colmajor(z["v"]), z["w"]
    dt = float(z["dt"])
    fc = mesh.l2_project(v, w)
    S0 = mesh.entropy(fc)
    plain = O.default_params(dt, m=5)
    v1, _ = mesh.predictor(plain, v, w, fc)
    _, st_plain, _ = mesh.step_anderson(plain, v1, v, w, S0)
    cl = O.default_params(dt, m=5, cl_theta=0.005 / dt, cl_kmax=32)
    x, st, res = mesh.step_anderson(cl, v1, v, w, S0)
    print(f"plain: {st_plain.iters} maps exit {st_plain.exit_reason} res {st_plain.residual:.2e}; "
          f"clusters ({st.n_clusters}, {st.cl_fail} failed solves): {st.iters} maps exit {st.exit_reason} "
          f"res {st.residual:.2e}")
    assert st_plain.exit_reason != 0
    assert st.exit_reason == 0 and st.n_clusters > 10
    tol = max(1e-12 * np.linalg.norm(x), 1e-10)
    # x = G(v*) for the last iterate v*; a root of the plain map to the tolerance
    assert np.linalg.norm(mesh.picard_map(plain, x, v, w, S0) - x) <= 20 * tol
    # momentum to roundoff, energy to the step's residual level
    P0, E0 = O.momentum_energy(v, w)
    P1, E1 = O.momentum_energy(x, w)
    assert max(abs(P1[0] - P0[0]), abs(P1[1] - P0[1])) <= 1e-13
    assert abs(E1 - E0) / E0 <= 1e-10
