"""Parity of every hook-level kernel of libcollop.so against the oracle.

Inputs are seeded markers fed as identical float64 bytes to both sides; the
entropy-gradient field G comes from the oracle's own projection chain (as in
scripts/check_metric_gpu.jl:28-30).  Errors are normwise, max|d| / max|ref|
(scripts/check_metric_gpu.jl:48-51).  Bars: pair sum 1e-12, projection
coefficients 1e-13 (FP64); the FP32-compensated pair sum 2e-6.
"""
import numpy as np
import pytest

import collop
import oracle as O
from conftest import relerr, uniform_bp

pytestmark = pytest.mark.gpu


def chain(ic, N, ne, lo=-6.0, hi=6.0, seed=42, jitter=0.0):
    sig = {"gaussian": (1.0, 1.0), "bimodal": (np.sqrt(0.5), np.sqrt(0.5)),
           "bimaxwellian": (np.sqrt(2.0), np.sqrt(0.5))}[ic]
    v, w = collop.markers(ic, N, seed, sigma1=sig[0], sigma2=sig[1], peak=2.0, weight_jitter=jitter)
    bp = uniform_bp(lo, hi, ne)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    L = m.mass_solve(m.compute_r(fc))
    G = m.compute_G(v, L)
    return v, w, bp, m, fc, L, G


@pytest.fixture(scope="module")
def c1(gpu_ctx):
    v, w, bp, m, fc, L, G = chain("gaussian", 4096, 16)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    return dict(v=v, w=w, m=m, fc=fc, L=L, G=G, ws=ws)


def test_tables_match_oracle(c1):
    C1, C2, bs1, bs2 = c1["m"].tables()
    np.testing.assert_array_equal(collop.mesh.flat_tables(c1["ws"].ext1d_1), C1)
    np.testing.assert_array_equal(c1["ws"].basis_start_1d_1, bs1)


def test_collision_f64_n4096(c1):
    ref = c1["m"].collision(c1["v"], c1["w"], c1["G"])
    out = np.zeros_like(ref)
    collop.compute_collision_gpu(c1["ws"], out, c1["v"], c1["w"], c1["G"])
    e = relerr(out, ref)
    print(f"collision f64 N=4096 relerr {e:.3e}")
    assert e <= 1e-12


@pytest.mark.parametrize("tile", [(1, 0, 2), (2, 0, 2), (1, 8, 1), (2, 16, 1), (1, 1, 2), (2, 64, 2)])
def test_collision_f64_tiles(c1, tile):
    ref = c1["m"].collision(c1["v"], c1["w"], c1["G"])
    collop.set_pair_tile(c1["ws"], *tile)
    try:
        out = np.zeros_like(ref)
        collop.compute_collision_gpu(c1["ws"], out, c1["v"], c1["w"], c1["G"])
    finally:
        collop.set_pair_tile(c1["ws"], 1, 0, 1)
        collop.set_pair_mode(c1["ws"], True)
    assert relerr(out, ref) <= 1e-12


def test_collision_f64_bimodal_n65536_subsample(gpu_ctx):
    v, w, bp, m, fc, L, G = chain("bimodal", 65536, 32)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    out = np.zeros((65536, 2), order="F")
    collop.compute_collision_gpu(ws, out, v, w, G)
    ref = O.collision_rows(m.box, v, w, G, 0, 3000)
    e = relerr(out[:3000], ref[:3000])
    print(f"collision f64 N=65536 (3000 targets) relerr {e:.3e}")
    assert e <= 1e-12


def test_collision_f32c(c1):
    ref = c1["m"].collision(c1["v"], c1["w"], c1["G"])
    out = np.zeros_like(ref)
    collop.compute_collision_gpu32(c1["ws"], out, c1["v"], c1["w"], c1["G"])
    e = relerr(out, ref)
    print(f"collision f32c N=4096 relerr {e:.3e}")
    assert e <= 2e-6
    # every pair term is exactly antisymmetric in FP32 and the weights are equal,
    # so the momentum sum cancels to the rounding of the compensated sums
    p = (c1["w"][:, None] * out).sum(axis=0)
    assert np.all(np.abs(p) <= 1e-9 * np.abs(out).max())


def test_collision_metric(c1):
    F_ref, A_ref = O.collision_metric(c1["m"].box, c1["v"], c1["w"], c1["G"])
    F = np.zeros_like(F_ref)
    A = np.zeros_like(A_ref)
    collop.compute_collision_metric_gpu(c1["ws"], F, A, c1["v"], c1["w"], c1["G"])
    assert relerr(F, F_ref) <= 1e-12
    assert relerr(A, A_ref) <= 1e-12


def test_collision_metric_f32c(c1):
    F_ref, A_ref = O.collision_metric(c1["m"].box, c1["v"], c1["w"], c1["G"])
    F = np.zeros_like(F_ref)
    A = np.zeros_like(A_ref)
    collop.compute_collision_metric_gpu(c1["ws"], F, A, c1["v"], c1["w"], c1["G"], fp32c=True)
    eF, eA = relerr(F, F_ref), relerr(A, A_ref)
    print(f"metric f32c N=4096 relerr F {eF:.3e} A {eA:.3e}")
    # F: pair terms vanish as g = G_a - G_g -> 0 for close pairs, so FP32 rounding
    # of the positions barely matters.  A = sum w U(d) weights close pairs by
    # 1 / |d| with no such factor: an FP32 position error of ~1e-7 is ~1e-4
    # relative at |d| ~ 1e-3.  A only preconditions the frozen-metric inner
    # solve of defect correction and cannot move its fixed point.
    assert eF <= 2e-6 and eA <= 1e-4


def test_l2_rhs_and_project(c1):
    m, ws = c1["m"], c1["ws"]
    rhs = np.zeros(ws.n_dofs)
    collop.l2_rhs_gpu(ws, rhs, c1["v"], c1["w"])
    assert relerr(rhs, m.l2_rhs(c1["v"], c1["w"])) <= 1e-13
    fc = np.zeros(ws.n_dofs)
    collop.l2_project_gpu(ws, fc, c1["v"], c1["w"])
    e = relerr(fc, c1["fc"])
    print(f"l2_project relerr {e:.3e}")
    assert e <= 1e-13


def test_mass_solve(c1):
    m, ws = c1["m"], c1["ws"]
    b = np.random.default_rng(3).standard_normal(ws.n_dofs)
    assert relerr(collop.mass_solve_gpu(ws, b), m.mass_solve(b)) <= 1e-13


@pytest.mark.parametrize("logsq", [False, True])
def test_quadrature(c1, logsq):
    m, ws = c1["m"], c1["ws"]
    r, S = collop.compute_r_and_entropy_gpu(ws, c1["fc"], use_logsq=logsq)
    assert relerr(r, m.compute_r(c1["fc"], logsq)) <= 1e-13
    assert abs(S - m.entropy(c1["fc"], logsq)) <= 1e-13 * abs(S)


def test_grad_gather(c1):
    G = np.zeros((4096, 2), order="F")
    collop.compute_G_gpu(c1["ws"], G, c1["v"], c1["L"])
    assert relerr(G, c1["G"]) <= 1e-13


def test_loggrad_gather(c1):
    g = np.zeros((4096, 2), order="F")
    collop.eval_loggrad_gpu(c1["ws"], g, c1["v"], c1["fc"])
    assert relerr(g, c1["m"].loggrad(c1["v"], c1["fc"])) <= 1e-13


def test_out_of_domain_and_coincident(gpu_ctx):
    bp = uniform_bp(-2.0, 2.0, 8)
    m = O.Mesh(bp, bp)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    rng = np.random.default_rng(11)
    v = np.asfortranarray(rng.uniform(-2.5, 2.5, (300, 2)))
    v[5] = v[6]                     # coincident pair: skipped
    v[7] = (-2.0, 0.3)              # exactly on the boundary: outside the open box
    v[8] = (2.0, 2.0)
    w = np.full(300, 1 / 300)
    G = np.asfortranarray(rng.standard_normal((300, 2)))
    ref = m.collision(v, w, G)
    out = np.zeros_like(ref)
    collop.compute_collision_gpu(ws, out, v, w, G)
    assert relerr(out, ref) <= 1e-12
    outside = (v[:, 0] <= -2) | (v[:, 0] >= 2) | (v[:, 1] <= -2) | (v[:, 1] >= 2)
    assert outside.sum() > 10
    assert np.all(out[outside] == 0.0)
    L = rng.standard_normal(ws.n_dofs)
    Gd = np.zeros((300, 2), order="F")
    collop.compute_G_gpu(ws, Gd, v, L)
    assert np.all(Gd[outside] == 0.0)
    assert relerr(Gd, m.compute_G(v, L)) <= 1e-13


def test_nonuniform_mesh(gpu_ctx):
    bp1 = np.array([-6.0, -5.0, -4.0, -3.0, -1.5, 0.0, 0.7, 2.0, 4.0, 5.5, 6.0])
    bp2 = np.array([-6.0, -4.5, -3.0, -1.0, 0.0, 1.0, 3.0, 4.5, 6.0])
    m = O.Mesh(bp1, bp2)
    ws = collop.Workspace(bp1, bp2, ctx=gpu_ctx)
    v, w = collop.markers("gaussian", 2000, 5, sigma1=1.5, sigma2=1.2)
    fc = np.zeros(ws.n_dofs)
    collop.l2_project_gpu(ws, fc, v, w)
    fr = m.l2_project(v, w)
    assert relerr(fc, fr) <= 1e-13
    r, S = collop.compute_r_and_entropy_gpu(ws, fr)
    assert relerr(r, m.compute_r(fr)) <= 1e-13
    L = m.mass_solve(m.compute_r(fr))
    G = np.zeros((2000, 2), order="F")
    collop.compute_G_gpu(ws, G, v, L)
    assert relerr(G, m.compute_G(v, L)) <= 1e-13


def test_empty_and_errors(gpu_ctx):
    bp = uniform_bp(-6, 6, 16)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    out = np.zeros((0, 2), order="F")
    collop.compute_collision_gpu(ws, out, np.zeros((0, 2)), np.zeros(0), np.zeros((0, 2)))
    fc = np.zeros(ws.n_dofs)
    collop.l2_project_gpu(ws, fc, np.zeros((0, 2)), np.zeros(0))
    assert np.all(fc == 0.0)
    with pytest.raises(collop.CollopError):
        collop.Workspace(bp, bp, p_deg=3, ctx=gpu_ctx)
    with pytest.raises(collop.CollopError) as ei:
        collop.eval_loggrad_gpu(ws, np.zeros((4, 2)), np.zeros((4, 2)), fc, fs_floor=-1.0)
    assert ei.value.status == collop._lib.ERR_INVALID


def test_deterministic(c1):
    a = np.zeros((4096, 2), order="F")
    b = np.zeros((4096, 2), order="F")
    collop.compute_collision_gpu(c1["ws"], a, c1["v"], c1["w"], c1["G"])
    collop.compute_collision_gpu(c1["ws"], b, c1["v"], c1["w"], c1["G"])
    np.testing.assert_array_equal(a, b)
