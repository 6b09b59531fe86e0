"""Stiff-cluster elimination on the device (csrc/cluster.hip) against the
oracle: on a state where close pairs have formed (the oracle's own Anderson
solve of the next step stagnates), the device step with clusters converges, it
finds the oracle's clusters, and its result is a root of the oracle's implicit
system; momentum and energy are conserved and entropy rises.  Then the same
check at the headline size: one step of the chosen path, of the Anderson path
with clusters and of the plain FP64 path at N = 65536 from the bimodal initial
state."""
import os

import numpy as np
import pytest

import collop
import oracle as O
from conftest import uniform_bp

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stiff_bimodal_n8192.npz")


def _check_root(m, v_new, v, w, dt):
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(dt, m=5)
    tol = max(1e-12 * np.linalg.norm(v_new), 1e-10)
    res = float(np.linalg.norm(m.picard_map(P, v_new, v, w, S0) - v_new))
    return res, tol, S0


@pytest.mark.parametrize("precision", [collop.PREC_F64, collop.PREC_F32C])
def test_stiff_step_converges_to_an_oracle_root(gpu_ctx, precision):
    z = np.load(GOLD)
    v, w, dt = np.asfortranarray(z["v"]), z["w"], float(z["dt"])
    bp = uniform_bp(-6, 6, 32)
    m = O.Mesh(bp, bp)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    theta = 0.005 / dt
    p = collop.config("bimodal_n65536", N_PARTICLES=v.shape[0], cl_theta=theta, cl_kmax=32)
    sim = collop.Simulation(ws, v, w)
    st = sim.run(p.step_params(precision=precision), 1)[0]
    x = sim.markers()
    n_oracle = len(m.clusters(v, w, theta, 32))
    res, tol, S0 = _check_root(m, x, v, w, dt)
    print(f"precision {precision}: {st['iters']} maps (f32 {st['iters_f32']}), exit {st['exit_reason']}, "
          f"clusters {st['n_clusters']} (oracle {n_oracle}), failed solves {st['cl_fail']}; "
          f"oracle map residual {res:.2e} (tol {tol:.2e})")
    assert st["n_clusters"] == n_oracle
    assert st["exit_reason"] == 0
    assert res <= 20 * tol
    P0, E0 = O.momentum_energy(v, w)
    P1, E1 = O.momentum_energy(x, w)
    assert max(abs(P1[0] - P0[0]), abs(P1[1] - P0[1])) <= 1e-13
    assert abs(E1 - E0) / E0 <= 1e-10
    assert st["entropy"] >= S0


def test_stiff_step_plain_solve_stalls(gpu_ctx):
    """The same step without the cluster elimination: the reference's Anderson
    solve stagnates (exit 1) or needs many more maps."""
    z = np.load(GOLD)
    v, w, dt = np.asfortranarray(z["v"]), z["w"], float(z["dt"])
    bp = uniform_bp(-6, 6, 32)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    p0 = collop.config("bimodal_n65536", N_PARTICLES=v.shape[0])
    p1 = collop.config("bimodal_n65536", N_PARTICLES=v.shape[0], cl_theta=0.005 / dt)
    a = collop.Simulation(ws, v, w).run(p0.step_params(), 1)[0]
    b = collop.Simulation(ws, v, w).run(p1.step_params(), 1)[0]
    print(f"plain: {a['iters']} maps exit {a['exit_reason']}; clusters: {b['iters']} maps exit {b['exit_reason']}")
    assert a["exit_reason"] != 0 or a["iters"] > 2 * b["iters"]
    assert b["exit_reason"] == 0


def test_chosen_step_at_n65536_is_an_oracle_root(gpu_ctx):
    """One step of the chosen path (FP32c defect correction with the
    hand-over, learned warm start) at N = 65536 on 32 x 32 cells, dt 0.02: the
    state is a root of the oracle's implicit system; and one step of the
    Anderson path with stiff clusters, and of the plain FP64 path, is too."""
    from collop import warmstart as W
    p = collop.chosen_config("bimodal_n65536")
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, ctx=gpu_ctx)
    W.load_packaged(ws)
    m = O.Mesh(p.bp1, p.bp2)
    for name, prm in (("chosen", p.step_params(warmstart=collop.WARM_MLP)),
                      ("anderson + clusters", collop.cluster_config("bimodal_n65536").step_params(
                          warmstart=collop.WARM_MLP)),
                      ("plain FP64", collop.config("bimodal_n65536").step_params())):
        sim = collop.Simulation(ws, v, w)
        st = sim.run(prm, 1)[0]
        x = sim.markers()
        res, tol, S0 = _check_root(m, x, v, w, p.DT)
        print(f"{name}: {st['iters']} maps, exit {st['exit_reason']}, oracle map residual {res:.2e} (tol {tol:.2e})")
        assert st["exit_reason"] == 0
        assert res <= 20 * tol
        assert st["entropy"] >= S0
