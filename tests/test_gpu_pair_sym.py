"""The symmetric pair sum (every unordered pair evaluated once and applied to
both markers) against the oracle's ordered sum (functions.jl:205-237,
:249-290), for every block shape, ragged N (partial super-rows and tiles),
markers outside the box and coincident pairs in different super-rows, and
against the ordered GPU kernel.  Bars as for the ordered kernel: FP64 1e-12,
FP32-compensated 2e-6 (metric A 1e-4, see test_gpu_kernels.py).
"""
import numpy as np
import pytest

import collop
import oracle as O
from conftest import relerr, uniform_bp

pytestmark = pytest.mark.gpu


def chain(ic, N, ne, seed=42):
    sig = {"gaussian": (1.0, 1.0), "bimodal": (np.sqrt(0.5), np.sqrt(0.5))}[ic]
    v, w = collop.markers(ic, N, seed, sigma1=sig[0], sigma2=sig[1], peak=2.0, weight_jitter=0.3)
    bp = uniform_bp(-6.0, 6.0, ne)
    m = O.Mesh(bp, bp)
    fc = m.l2_project(v, w)
    G = m.compute_G(v, m.mass_solve(m.compute_r(fc)))
    return v, w, bp, m, G


def all_four(ws, v, w, G):
    N = v.shape[0]
    F64 = np.zeros((N, 2), order="F")
    F32 = np.zeros((N, 2), order="F")
    MF = np.zeros((N, 2), order="F")
    MA = np.zeros((N, 3), order="F")
    MF32 = np.zeros((N, 2), order="F")
    MA32 = np.zeros((N, 3), order="F")
    collop.compute_collision_gpu(ws, F64, v, w, G)
    collop.compute_collision_gpu32(ws, F32, v, w, G)
    collop.compute_collision_metric_gpu(ws, MF, MA, v, w, G)
    collop.compute_collision_metric_gpu(ws, MF32, MA32, v, w, G, fp32c=True)
    return F64, F32, MF, MA, MF32, MA32


@pytest.fixture(autouse=True)
def default_pair_mode(gpu_ctx):
    """The pair form is a setting of the shared context: restore the default."""
    yield
    gpu_ctx.check(gpu_ctx.lib.collop_set_pair_mode(gpu_ctx.h, collop._lib.PAIR_SYMMETRIC, 0))


@pytest.fixture(scope="module")
def sym_ws(gpu_ctx):
    return collop.Workspace(uniform_bp(-6.0, 6.0, 16), uniform_bp(-6.0, 6.0, 16), ctx=gpu_ctx)


@pytest.mark.parametrize("tiles", [2, 4, 8])
@pytest.mark.parametrize("N", [1, 2, 129, 1000, 4096, 5003])
def test_symmetric_matches_oracle(sym_ws, tiles, N):
    v, w, bp, m, G = chain("gaussian", N, 16, seed=N)
    collop.set_pair_mode(sym_ws, True, tiles)
    F64, F32, MF, MA, MF32, MA32 = all_four(sym_ws, v, w, G)
    ref = m.collision(v, w, G)
    Fr, Ar = O.collision_metric(m.box, v, w, G)
    e = [relerr(F64, ref), relerr(F32, ref), relerr(MF, Fr), relerr(MA, Ar), relerr(MF32, Fr), relerr(MA32, Ar)]
    print(f"tiles {tiles} N {N}: f64 {e[0]:.1e} f32c {e[1]:.1e} metric F {e[2]:.1e} A {e[3]:.1e} "
          f"f32c metric F {e[4]:.1e} A {e[5]:.1e}")
    if N == 1:  # a single marker has no partner
        assert np.all(F64 == 0.0) and np.all(MA == 0.0)
        return
    assert e[0] <= 1e-12 and e[2] <= 1e-12 and e[3] <= 1e-12
    assert e[1] <= 2e-6 and e[4] <= 2e-6 and e[5] <= 1e-4


def test_symmetric_out_of_domain_and_coincident(gpu_ctx):
    bp = uniform_bp(-2.0, 2.0, 8)
    m = O.Mesh(bp, bp)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    collop.set_pair_mode(ws, True, 2)  # super-rows of 256: three of them at N = 700
    rng = np.random.default_rng(12)
    N = 700
    v = np.asfortranarray(rng.uniform(-2.5, 2.5, (N, 2)))
    v[5] = v[600]                   # coincident pair in different super-rows: skipped
    v[7] = v[8]                     # and in the same tile
    v[9] = (-2.0, 0.3)              # on the boundary: outside the open box
    w = rng.uniform(0.5, 1.5, N) / N
    G = np.asfortranarray(rng.standard_normal((N, 2)))
    ref = m.collision(v, w, G)
    Fr, Ar = O.collision_metric(m.box, v, w, G)
    F64, F32, MF, MA, MF32, MA32 = all_four(ws, v, w, G)
    collop.set_pair_mode(ws, False)
    o = all_four(ws, v, w, G)
    assert relerr(F64, ref) <= 1e-12 and relerr(MF, Fr) <= 1e-12 and relerr(MA, Ar) <= 1e-12
    # G here is noise, not a projected gradient: G_a - G_g stays O(1) for close
    # pairs, so the FP32 rounding of the positions (identical in both kernels:
    # d is formed the same way) is not cancelled and both FP32c kernels sit
    # ~3e-6 from the FP64 oracle.  Against each other they differ by summation
    # order only.
    e32 = [relerr(F32, ref), relerr(F32, o[1]), relerr(MF32, o[4]), relerr(MA32, o[5])]
    print("f32c: vs oracle %.1e, vs ordered kernel F %.1e metric F %.1e A %.1e" % tuple(e32))
    assert e32[0] <= 1e-5 and relerr(o[1], ref) <= 1e-5
    assert e32[1] <= 1e-6 and e32[2] <= 1e-6 and e32[3] <= 1e-6
    outside = (v[:, 0] <= -2) | (v[:, 0] >= 2) | (v[:, 1] <= -2) | (v[:, 1] >= 2)
    assert outside.sum() > 20
    for X in (F64, F32, MF, MA, MF32, MA32):
        assert np.all(X[outside] == 0.0)


def test_symmetric_vs_ordered_bimodal_n65536(gpu_ctx):
    v, w, bp, m, G = chain("bimodal", 65536, 32)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    collop.set_pair_mode(ws, False)
    o = all_four(ws, v, w, G)
    collop.set_pair_mode(ws, True)
    s = all_four(ws, v, w, G)
    ref = O.collision_rows(m.box, v, w, G, 0, 2000)
    e_sym = relerr(s[0][:2000], ref[:2000])
    e_ord = relerr(o[0][:2000], ref[:2000])
    d = [relerr(a, b) for a, b in zip(s, o)]
    print(f"N=65536: oracle rows 0..2000 sym {e_sym:.2e} ordered {e_ord:.2e}; sym vs ordered "
          + " ".join(f"{x:.1e}" for x in d))
    assert e_sym <= 1e-12 and e_ord <= 1e-12
    assert d[0] <= 1e-13 and d[2] <= 1e-13 and d[3] <= 1e-13
    assert d[1] <= 2e-6 and d[4] <= 2e-6 and d[5] <= 1e-4
    # momentum: each unordered pair adds w_g (w_a t) - w_a (w_g t) = O(rounding) to sum w F
    scale = np.abs(w[:, None] * s[0]).sum(axis=0)
    p_sym = np.abs((w[:, None] * s[0]).sum(axis=0)) / scale
    p_ord = np.abs((w[:, None] * o[0]).sum(axis=0)) / scale
    print(f"relative momentum of the pair sum: sym {p_sym} ordered {p_ord}")
    assert np.all(p_sym <= 1e-12)


def test_symmetric_deterministic_and_fallback(gpu_ctx):
    v, w, bp, m, G = chain("gaussian", 66000, 16, seed=3)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    collop.set_pair_mode(ws, True, 8)
    a = np.zeros((66000, 2), order="F")
    b = np.zeros((66000, 2), order="F")
    collop.compute_collision_gpu(ws, a, v, w, G)
    collop.compute_collision_gpu(ws, b, v, w, G)
    np.testing.assert_array_equal(a, b)
    # 66000 markers take 258 super-rows of 256 at tiles = 2: still symmetric
    collop.set_pair_mode(ws, True, 2)
    assert collop.pair_mode(ws, 66000) == (True, 2, 258)
    c = np.zeros((66000, 2), order="F")
    collop.compute_collision_gpu(ws, c, v, w, G)
    ref = O.collision_rows(m.box, v, w, G, 0, 1500)
    assert relerr(c[:1500], ref[:1500]) <= 1e-12
    assert relerr(a[:1500], ref[:1500]) <= 1e-12
    assert relerr(a, c) <= 1e-13
    with pytest.raises(collop.CollopError):
        collop.set_pair_mode(ws, True, 3)


def test_symmetric_beyond_256_slots_and_fallback(gpu_ctx):
    """300000 markers take 293 super-rows of 1024 (more than 256 partial
    slots): the symmetric sum against the ordered one and the oracle; 140000
    markers at 2 waves per block would need 547 slots (more than 512): the
    ordered sum runs."""
    v, w, bp, m, G = chain("gaussian", 300000, 16, seed=5)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    collop.set_pair_mode(ws, True)
    assert collop.pair_mode(ws, 300000) == (True, 8, 293)
    s = np.zeros((300000, 2), order="F")
    collop.compute_collision_gpu(ws, s, v, w, G)
    collop.set_pair_mode(ws, False)
    o = np.zeros((300000, 2), order="F")
    collop.compute_collision_gpu(ws, o, v, w, G)
    ref = O.collision_rows(m.box, v, w, G, 0, 500)
    ref2 = O.collision_rows(m.box, v, w, G, 299500, 300000)
    e = max(relerr(s[:500], ref[:500]), relerr(s[299500:], ref2[299500:]))
    print(f"N=300000: symmetric vs oracle (1000 rows) {e:.2e}, vs ordered {relerr(s, o):.2e}")
    assert e <= 1e-12 and relerr(s, o) <= 1e-13
    collop.set_pair_mode(ws, True, 2)
    assert collop.pair_mode(ws, 140000)[:2] == (False, 0)


def test_pair_mode_choice(sym_ws):
    assert collop.pair_mode(sym_ws, 1000) == (True, 2, 4)
    assert collop.pair_mode(sym_ws, 65536) == (True, 4, 128)
    assert collop.pair_mode(sym_ws, 262144) == (True, 8, 256)
    assert collop.pair_mode(sym_ws, 300000) == (True, 8, 293)
    assert collop.pair_mode(sym_ws, 600000)[:2] == (False, 0)  # beyond 512 slots: ordered
    collop.set_pair_mode(sym_ws, True, 2)
    assert collop.pair_mode(sym_ws, 65536) == (True, 2, 256)
    collop.set_pair_mode(sym_ws, False)
    assert collop.pair_mode(sym_ws, 65536) == (False, 0, 32)
    collop.set_pair_tile(sym_ws, 2, 16, 1)
    assert collop.pair_mode(sym_ws, 65536) == (False, 0, 16)
    collop.set_pair_tile(sym_ws, 1, 0, 1)


@pytest.mark.parametrize("tiles", [2, 4, 8])
@pytest.mark.parametrize("N", [1000, 5003])
def test_equal_weight_hooks_are_bitwise_the_general_form(sym_ws, tiles, N):
    """The collision hooks with every weight 2^-12 (5 % of the markers outside
    the box, coincident pairs in different super-rows, which send their blocks
    through the guarded rerun): the equal-weight forms (plain FP64, plain and
    metric FP32-compensated) give the general forms' bits, and stay within the
    oracle's bars.  Then jittered weights after a state of equal weights: the
    hooks judge the weights of their own call, not the state's."""
    v, _, bp, m, _ = chain("gaussian", N, 16, seed=N)
    rng = np.random.default_rng(N)
    out = rng.choice(N, N // 20, replace=False)
    v[out] *= 6.5 / np.max(np.abs(v[out]), axis=1, keepdims=True)
    for i, j in ((5, 600), (7, 8), (40, N - 3)):
        v[i] = v[j]
    w = np.full(N, 2.0 ** -12)
    G = m.compute_G(v, m.mass_solve(m.compute_r(m.l2_project(v, w))))
    collop.set_pair_mode(sym_ws, True, tiles)
    res = {}
    for on in (True, False):
        collop.set_weight_path(sym_ws, on)
        res[on] = all_four(sym_ws, v, w, G)
    collop.set_weight_path(sym_ws, True)
    for a, b in zip(res[True], res[False]):
        np.testing.assert_array_equal(a, b)
    F64, F32, MF, MA, MF32, MA32 = res[True]
    ref = m.collision(v, w, G)
    refF, refA = O.collision_metric(m.box, v, w, G)
    print(f"N={N} R={tiles}: vs oracle {relerr(F64, ref):.1e} {relerr(F32, ref):.1e} "
          f"{relerr(MF32, refF):.1e} {relerr(MA32, refA):.1e}")
    assert relerr(F64, ref) <= 1e-12 and relerr(F32, ref) <= 2e-6
    assert relerr(MF32, refF) <= 2e-6 and relerr(MA32, refA) <= 1e-4
    sim = collop.Simulation(sym_ws, v, w)  # a state of equal power-of-two weights
    wj = w * (1.0 + 0.3 * rng.random(N))
    Gj = m.compute_G(v, m.mass_solve(m.compute_r(m.l2_project(v, wj))))
    jit = all_four(sym_ws, v, wj, Gj)
    collop.set_weight_path(sym_ws, False)
    gen = all_four(sym_ws, v, wj, Gj)
    collop.set_weight_path(sym_ws, True)
    del sim
    for a, b in zip(jit, gen):
        np.testing.assert_array_equal(a, b)
    assert relerr(jit[0], m.collision(v, wj, Gj)) <= 1e-12


@pytest.mark.parametrize("N,precision", [(4096, "f64"), (4096, "f32c"), (65536, "f32c"), (5003, "f64"),
                                         (5003, "f32c")])
def test_equal_weight_form_is_bitwise_the_general_form(gpu_ctx, N, precision):
    """The equal-weight form of the symmetric sums (every marker w = 2^-k: 1/N,
    or 2^-12 at N = 5003, whose last super-row is partial) against the general
    form on the same steps: the same bits, with 3 % of the markers outside the
    box (their pairs vanish through the far point instead of the mask).  Two
    steps of the chosen path's solver, FP64 outer and Anderson maps included."""
    p = collop.chosen_config("bimodal_n65536", N_PARTICLES=N)
    v, w = collop.initial_markers(p)
    if np.frexp(w[0])[0] != 0.5:
        w = np.full(N, 2.0 ** -12)
    rng = np.random.default_rng(5)
    out = rng.choice(N, N * 3 // 100, replace=False)
    v[out] *= 6.5 / np.max(np.abs(v[out]), axis=1, keepdims=True)  # pushed onto |v|_inf = 6.5
    assert np.all(w == w[0]) and np.frexp(w[0])[0] == 0.5
    ws = collop.build_workspace(p, ctx=gpu_ctx)
    prm = p.step_params(precision=collop.PREC_F64 if precision == "f64" else collop.PREC_F32C,
                        warmstart=collop.WARM_EULER)
    res = {}
    for on in (True, False):
        collop.set_weight_path(ws, on)
        sim = collop.Simulation(ws, v, w)
        st = sim.run(prm, 2)
        res[on] = (sim.markers(), [(s["iters"], s["iters_f32"], s["inner_maps"], s["residual"]) for s in st])
    collop.set_weight_path(ws, True)
    print(f"N={N} {precision}: maps {res[True][1]}")
    assert res[True][1] == res[False][1]
    np.testing.assert_array_equal(res[True][0], res[False][0])
    np.testing.assert_array_equal(res[True][0][out], res[False][0][out])
