"""The committed golden fixtures (tests/golden/make_golden.py): the oracle
must keep reproducing them (CPU), and the HIP library must match them (GPU).
Loaded with numpy.load(allow_pickle=False)."""
import os

import numpy as np
import pytest

import collop
import oracle as O
from conftest import relerr

G_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G_DIR, name), allow_pickle=False)


def test_oracle_reproduces_pair_fixture():
    d = load("golden_pair.npz")
    out = O.collision(d["box"], d["v"], d["w"], d["G"])
    np.testing.assert_array_equal(out, d["dot_v"])
    assert relerr(out, d["dot_v_ld"]) < 1e-14


def test_oracle_reproduces_fem_fixture():
    d = load("golden_fem.npz")
    m = O.Mesh(d["bp1"], d["bp2"])
    np.testing.assert_allclose(m.l2_rhs(d["v"], d["w"]), d["rhs"], rtol=0, atol=1e-16)
    fc = m.l2_project(d["v"], d["w"])
    assert relerr(fc, d["fc"]) < 1e-14
    assert relerr(m.compute_r(d["fc"]), d["r"]) < 1e-14
    assert abs(m.entropy(d["fc"]) - float(d["S"])) < 1e-14
    assert relerr(m.compute_G(d["v"], d["L"]), d["G"]) < 1e-14
    assert relerr(m.loggrad(d["v"], d["fc"]), d["g"]) < 1e-14


def test_oracle_reproduces_step_fixture():
    d = load("golden_step.npz")
    m = O.Mesh(d["bp"], d["bp"])
    fc = m.l2_project(d["v"], d["w"])
    v1, fc1, st = m.time_step(O.default_params(float(d["dt"]), m=5), d["v"], d["w"], fc, m.entropy(fc))
    assert st.iters == int(d["iters"])
    assert np.max(np.abs(v1 - d["v_new"])) < 1e-12


@pytest.mark.gpu
def test_gpu_matches_fixtures(gpu_ctx):
    d = load("golden_pair.npz")
    bp = np.array([-6.0, 6.0])
    bpm = np.linspace(d["box"][0], d["box"][1], 9)
    ws = collop.Workspace(bpm, bpm, ctx=gpu_ctx)
    out = np.zeros((96, 2), order="F")
    collop.compute_collision_gpu(ws, out, d["v"], d["w"], d["G"])
    assert relerr(out, d["dot_v_ld"]) <= 1e-12
    del bp
    f = load("golden_fem.npz")
    ws = collop.Workspace(f["bp1"], f["bp2"], ctx=gpu_ctx)
    fc = np.zeros(ws.n_dofs)
    collop.l2_project_gpu(ws, fc, f["v"], f["w"])
    assert relerr(fc, f["fc"]) <= 1e-13
    r, S = collop.compute_r_and_entropy_gpu(ws, f["fc"])
    assert relerr(r, f["r"]) <= 1e-13 and abs(S - float(f["S"])) <= 1e-13
    G = np.zeros((700, 2), order="F")
    collop.compute_G_gpu(ws, G, f["v"], f["L"])
    assert relerr(G, f["G"]) <= 1e-13
    s = load("golden_step.npz")
    ws = collop.Workspace(s["bp"], s["bp"], ctx=gpu_ctx)
    sim = collop.Simulation(ws, s["v"], s["w"])
    st = sim.run(collop.config("gaussian_n4096", DT=float(s["dt"])).step_params(), 1)[0]
    assert abs(st["iters"] - int(s["iters"])) <= 1
    assert np.max(np.abs(sim.markers() - s["v_new"])) <= 1e-9
