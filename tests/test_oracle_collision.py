"""Pins the oracle's pair sum (functions.jl:205-290) without trusting its own
formula: a two-particle closed form, a long-double brute force written as a
matrix product U(d) g, and the invariants of the Landau operator (momentum and
energy antisymmetry, G -> G + c, translation, rotation, scaling U(l d) =
U(d) / l), plus the domain mask and the coincident-pair guard.
"""
import numpy as np
import pytest

import oracle as O

BOX = np.array([-6.0, 6.0, -6.0, 6.0])


def brute_longdouble(v, w, G, box=BOX):
    """sum_a w_a U(d)(G_a - G_g) with U(d) = (I - d d^T/|d|^2)/|d| as an explicit 2x2 matrix."""
    ld = np.longdouble
    v = np.asarray(v, dtype=ld)
    G = np.asarray(G, dtype=ld)
    w = np.asarray(w, dtype=ld)
    N = len(v)
    inside = (v[:, 0] > box[0]) & (v[:, 0] < box[1]) & (v[:, 1] > box[2]) & (v[:, 1] < box[3])
    out = np.zeros((N, 2), dtype=ld)
    for g in range(N):
        if not inside[g]:
            continue
        d = v[g] - v                       # (N, 2)
        q = (d * d).sum(axis=1)
        ok = inside & (q >= ld(1e-24))
        ok[g] = False
        dd, qq = d[ok], q[ok]
        gg = G[ok] - G[g]
        r = np.sqrt(qq)
        # U(d) g = (g - d (d.g)/|d|^2) / |d|
        U_g = (gg - dd * ((dd * gg).sum(axis=1) / qq)[:, None]) / r[:, None]
        out[g] = (w[ok][:, None] * U_g).sum(axis=0)
    return out


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)))
                 / np.max(np.abs(np.asarray(b, dtype=np.float64))))


def sample(N, seed=0, spread=1.5):
    rng = np.random.default_rng(seed)
    v = O.colmajor(rng.normal(0, spread, (N, 2)))
    G = O.colmajor(rng.normal(0, 1, (N, 2)))
    w = rng.uniform(0.5, 1.5, N)
    w /= w.sum()
    return v, w, G


def test_two_particle_closed_form():
    v = O.colmajor([[0.3, -0.2], [1.1, 0.5]])
    G = O.colmajor([[0.7, 0.1], [-0.4, 2.0]])
    w = np.array([0.25, 0.75])
    out = O.collision(BOX, v, w, G)
    d = v[0] - v[1]
    g = G[1] - G[0]
    r = np.hypot(*d)
    exp0 = w[1] * (g - d * (d @ g) / r ** 2) / r
    exp1 = w[0] * (-g - (-d) * ((-d) @ (-g)) / r ** 2) / r
    np.testing.assert_allclose(out[0], exp0, rtol=1e-15)
    np.testing.assert_allclose(out[1], exp1, rtol=1e-15)


@pytest.mark.parametrize("N,seed", [(64, 1), (257, 2)])
def test_brute_force_longdouble(N, seed):
    v, w, G = sample(N, seed)
    v[3] = (7.0, 0.0)        # outside: zero and skipped
    v[4] = (6.0, 1.0)        # on the boundary: outside the open box
    out = O.collision(BOX, v, w, G)
    ref = brute_longdouble(v, w, G)
    assert rel(out, ref) < 1e-14
    assert np.all(out[3] == 0.0) and np.all(out[4] == 0.0)


def test_antisymmetry_conserves_momentum_and_energy():
    v, w, G = sample(400, 3)
    out = O.collision(BOX, v, w, G)
    scale = np.abs(w[:, None] * out).sum()
    assert np.all(np.abs((w[:, None] * out).sum(axis=0)) < 1e-14 * scale)
    assert abs(np.sum(w * (v[:, 0] * out[:, 0] + v[:, 1] * out[:, 1]))) < 1e-14 * scale * np.abs(v).max()


def test_invariances():
    v, w, G = sample(300, 4, spread=1.0)
    base = O.collision(BOX, v, w, G)
    # G -> G + const
    assert rel(O.collision(BOX, v, w, G + np.array([3.0, -1.5])), base) < 1e-13
    # translation (box moves with the markers)
    c = np.array([0.75, -0.5])
    assert rel(O.collision(BOX + np.array([c[0], c[0], c[1], c[1]]), v + c, w, G), base) < 1e-13
    # rotation by 90 degrees: (x, y) -> (-y, x) for v and G, and for the result
    R = np.array([[0.0, -1.0], [1.0, 0.0]])
    out = O.collision(BOX, O.colmajor(v @ R.T), w, O.colmajor(G @ R.T))
    assert rel(out, base @ R.T) < 1e-14
    # scaling: U(l d) = U(d) / l
    lam = 0.5
    assert rel(O.collision(BOX * lam, O.colmajor(v * lam), w, G), base / lam) < 1e-14


def test_masks():
    v, w, G = sample(200, 5)
    v[10] = v[11]                       # coincident pair: skipped by the 1e-24 guard
    G[10] = (5.0, 5.0)
    out = O.collision(BOX, v, w, G)
    ref = brute_longdouble(v, w, G)
    assert rel(out, ref) < 1e-14
    # an out-of-domain source is skipped exactly: removing it changes nothing
    v2 = v.copy()
    v2[20] = (9.0, 9.0)
    keep = np.ones(200, bool)
    keep[20] = False
    a = O.collision(BOX, v2, w, G)[keep]
    b = O.collision(BOX, O.colmajor(v2[keep]), w[keep], O.colmajor(G[keep]))
    np.testing.assert_array_equal(a, b)


def test_metric_consistent_with_pair_sum():
    v, w, G = sample(250, 6)
    F, A = O.collision_metric(BOX, v, w, G)
    np.testing.assert_array_equal(F, O.collision(BOX, v, w, G))
    # A_g = sum_a w_a U(d): F_g = B_g - A_g G_g with B_g = sum_a w_a U(d) G_a (solver.jl:539-543)
    ld = np.longdouble
    for g in (0, 17, 101):
        d = v[g] - v
        q = (d * d).sum(axis=1)
        ok = q > 0
        U = np.zeros((2, 2))
        B = np.zeros(2)
        for a in np.nonzero(ok)[0]:
            r = np.sqrt(ld(q[a]))
            Ua = (np.eye(2) - np.outer(d[a], d[a]) / ld(q[a])) / r
            U += w[a] * Ua.astype(np.float64)
            B += w[a] * (Ua @ G[a]).astype(np.float64)
        np.testing.assert_allclose(A[g], [U[0, 0], U[0, 1], U[1, 1]], rtol=1e-12)
        np.testing.assert_allclose(F[g], B - U @ G[g], rtol=1e-11, atol=1e-14)
        assert U[0, 0] > 0 and U[0, 0] * U[1, 1] - U[0, 1] ** 2 > 0   # positive definite


def test_rows_and_threads_agree():
    v, w, G = sample(500, 7)
    full = O.collision(BOX, v, w, G, nthreads=1)
    np.testing.assert_array_equal(full, O.collision(BOX, v, w, G, nthreads=4))
    part = O.collision_rows(BOX, v, w, G, 100, 230)
    np.testing.assert_array_equal(part[100:230], full[100:230])
