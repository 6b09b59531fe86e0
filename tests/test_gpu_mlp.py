"""The warm-start MLP kernels (BF16 MFMA, FP32 accumulate) against the numpy
reference oracle/mlp_ref.py, which rounds the same operands to bfloat16:
forward outputs, loss, and the full parameter gradient; plus a few Adam steps
that must lower the loss, and the recording / warm-start plumbing of a step."""
import numpy as np
import pytest

import collop
import mlp_ref as M
import oracle as O
from conftest import uniform_bp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wsm(gpu_ctx):
    bp = uniform_bp(-6, 6, 16)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    return ws, collop.WarmStart(ws)


def records(n, seed=0):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 8), np.float32)
    rec[:, :6] = rng.normal([0.5, -0.2, 1.0, 2.0, 0.1, 0.3], [1.5, 1.0, 3.0, 2.0, 0.05, 0.2], (n, 6))
    rec[:, 6:] = rng.normal(0, 0.05, (n, 2))
    return rec


def normalised(rec, mu, sd, os_):
    x = ((rec[:, :6] - mu) * (np.float32(1.0) / sd)).astype(np.float32)
    t = (rec[:, 6:] / os_).astype(np.float32)
    return x, t


def test_forward_loss_grad_match_numpy(wsm):
    ws, w = wsm
    p = M.init(3)
    # larger weights so activations leave GELU's linear range
    p[:384] *= 3.0
    rec = records(1000)
    mu = rec[:, :6].mean(0).astype(np.float32)
    sd = rec[:, :6].std(0).astype(np.float32)
    os_ = np.array([0.05, 0.04], np.float32)
    w.set_params(p)
    w.set_norm(mu, sd, os_)
    x, t = normalised(rec, mu, sd, os_)
    y, _, _ = w.eval(rec, 0)
    y_ref, _ = M.forward(p, x)
    err = np.max(np.abs(y - y_ref)) / np.max(np.abs(y_ref))
    print(f"forward rel err {err:.2e}")
    assert err < 2e-3
    _, loss, g = w.eval(rec, 2)
    loss_ref, g_ref = M.loss_and_grad(p, x, t)
    print(f"loss {loss:.6e} ref {loss_ref:.6e}; grad rel err {np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref)):.2e}")
    assert abs(loss - loss_ref) < 2e-3 * loss_ref
    for name, sl in [("W1", slice(0, 384)), ("b1", slice(384, 448)), ("W2", slice(448, 4544)),
                     ("b2", slice(4544, 4608)), ("W3", slice(4608, 4736)), ("b3", slice(4736, 4738))]:
        e = np.max(np.abs(g[sl] - g_ref[sl])) / np.max(np.abs(g_ref[sl]))
        assert e < 2e-2, (name, e)
    # the exact (unrounded float64) gradient of the restatement, itself pinned by
    # finite differences (tests/test_oracle_mlp.py), points the same way
    _, g64 = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), emulate_bf16=False)
    assert np.dot(g, g64) / (np.linalg.norm(g) * np.linalg.norm(g64)) > 0.995


def test_adam_lowers_loss(wsm):
    ws, w = wsm
    sim = collop.Simulation(ws, *collop.markers("gaussian", 4096, 5))
    w.dataset_alloc(4 * 4096)
    prm = collop.config("gaussian_n4096", DT=0.01).step_params()
    prm.record_training = 1
    sim.run(prm, 4)
    assert w.dataset_size() == 4 * 4096
    d = w.dataset_read(0, 4096)
    assert np.all(np.isfinite(d)) and np.abs(d[:, 6:]).max() > 0
    w.set_params(M.init(0))
    w.fit_norm()
    losses = w.train(epochs=3, batch=1024, lr=1e-3, seed=1)
    print("losses", losses)
    assert losses[-1] < losses[0]
    # warm start runs and the step still converges to the solver tolerance
    prm.record_training = 0
    prm.warmstart = collop.WARM_MLP
    st = sim.run(prm, 1)[0]
    assert st["exit_reason"] == 0


def test_device_adam_matches_oracle(gpu_ctx):
    """Three Adam steps of collop_mlp_train on a one-tile dataset (16 samples,
    batch 16: one step per epoch) against mlp_ref.adam fed the device's own
    gradient of that tile (collop_mlp_eval), so that the update rule is what is
    compared: moments, bias corrections and the parameter update."""
    bp = uniform_bp(-6, 6, 16)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    w = collop.WarmStart(ws)
    sim = collop.Simulation(ws, *collop.markers("gaussian", 16, 3))
    w.dataset_alloc(16)
    prm = collop.config("gaussian_n4096", DT=0.01).step_params()
    prm.record_training = 1
    sim.run(prm, 1)
    assert w.dataset_size() == 16
    rec = w.dataset_read(0, 16)
    p = M.init(4)
    w.set_params(p)
    w.fit_norm()
    m = np.zeros(M.NPARAM, np.float32)
    v = np.zeros(M.NPARAM, np.float32)
    lr = 1e-3
    for step in range(1, 4):
        _, _, g = w.eval(rec, 2)
        p_ref, m, v = M.adam(p, g.astype(np.float32), m, v, step, lr=lr)
        w.train(epochs=1, batch=16, lr=lr, seed=step)
        p_dev = w.params()
        err = float(np.max(np.abs(p_dev - p_ref)))
        moved = float(np.max(np.abs(p_ref - p)))
        print(f"Adam step {step}: max |device - oracle| {err:.2e} (step size {moved:.2e})")
        assert moved > 0.5 * lr
        assert err <= 1e-3 * lr
        p = p_dev


def test_trained_warm_start_gain(gpu_ctx):
    """The learned warm start end to end at a small size: record solver steps on
    the device, fit the normalisation, train with the library's own Adam, then
    step a held-out seed cold (Euler) and warm (Euler + MLP).  The network must
    shrink the predictor's error -- the first map's residual r0 -- and both
    solves must reach the same root (only the iteration changes)."""
    from collop import warmstart as W
    p = collop.config("bimodal_n65536", N_PARTICLES=4096, DT=0.005)
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    runs = W.training_runs(3, 4096, seed0=300)
    n = W.record(ws, runs, 40, p.step_params())
    assert n == 3 * 40 * 4096
    ml = W.WarmStart(ws)
    ml.set_params(W.glorot_init(0))
    ml.fit_norm()
    losses = ml.train(30, batch=4096, lr=1e-3, seed=2)
    print("loss per epoch", [round(x, 4) for x in losses[::5]])
    assert losses[-1] < losses[0]
    v, w = collop.markers("bimodal", 4096, 777, sigma1=0.5 ** 0.5, sigma2=0.5 ** 0.5, peak=2.0)
    cold = W.evaluate(ws, v, w, p.step_params(warmstart=collop.WARM_EULER), 5)
    warm = W.evaluate(ws, v, w, p.step_params(warmstart=collop.WARM_MLP), 5)
    print(f"held-out: r0 cold {cold['mean_r0']:.3e} warm {warm['mean_r0']:.3e} "
          f"(ratio {warm['mean_r0'] / cold['mean_r0']:.3f}); maps cold {cold['mean_maps']:.1f} warm {warm['mean_maps']:.1f}")
    assert cold["all_converged"] and warm["all_converged"]
    assert warm["mean_r0"] < cold["mean_r0"]
    # one step from the same state: both results are roots of the oracle's
    # implicit system to the solver tolerance.  They need not be the same
    # markers: this system (N = 4096, dt = 0.005) has states ~1e-5 apart whose
    # residuals both meet the tolerance (several roots, or a nearly neutral
    # direction of the map), and which one the solver reaches depends on the
    # initial guess
    c1 = W.evaluate(ws, v, w, p.step_params(warmstart=collop.WARM_EULER), 1)
    w1 = W.evaluate(ws, v, w, p.step_params(warmstart=collop.WARM_MLP), 1)
    assert c1["all_converged"] and w1["all_converged"]
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(p.DT, m=5, nthreads=0)
    for res in (c1, w1):
        x = res["markers"]
        tol = max(1e-12 * np.linalg.norm(x), 1e-10)
        assert np.linalg.norm(m.picard_map(P, x, v, w, S0) - x) <= 20 * tol
    diff = np.max(np.abs(w1["markers"] - c1["markers"]), axis=1)
    print(f"one step: max |dv| {diff.max():.2e}, markers beyond 1e-8: {int((diff > 1e-8).sum())}")


def test_packaged_model_loads_and_steps(gpu_ctx):
    from collop import warmstart as W
    p = collop.chosen_config("bimodal_n65536", N_PARTICLES=8192)
    ws = collop.build_workspace(p, ctx=gpu_ctx)
    ml = W.load_packaged(ws)
    assert ml.params().shape == (W.NPARAM,) and np.all(np.isfinite(ml.params()))
    v, w = collop.initial_markers(p)
    sim = collop.Simulation(ws, v, w)
    st = sim.run(p.step_params(warmstart=collop.WARM_MLP), 2)
    assert all(s["exit_reason"] == 0 for s in st)


def test_warm_start_at_n65536(gpu_ctx):
    """The learned warm start at the workload's size: the packaged model on a
    held-out bimodal seed (999), N = 65536 on 32 x 32 cells, dt 0.02, two steps
    of the chosen path with and without it.  Each state is a root of the
    oracle's implicit system, the model lowers the first residual r0 below the
    Euler predictor's, and the fixed point does not depend on the predictor.
    Then a short recording and training pass at the same size: the loss is
    finite and falls."""
    from collop import warmstart as W
    p = collop.chosen_config("bimodal_n65536")
    ws = collop.build_workspace(p, ctx=gpu_ctx)
    W.load_packaged(ws)
    v, w = collop.markers("bimodal", 65536, 999, sigma1=0.5 ** 0.5, sigma2=0.5 ** 0.5, peak=2.0)
    m = O.Mesh(p.bp1, p.bp2)
    P = O.default_params(p.DT, m=5)
    r0, last = {}, {}
    for k in ("euler", "mlp"):  # one simulation owns the context's state at a time
        sim = collop.Simulation(ws, v, w)
        prev = v
        r0[k] = []
        for step in range(2):
            st = sim.run(p.step_params(warmstart=collop.WARM_MLP if k == "mlp" else collop.WARM_EULER), 1)[0]
            x = sim.markers()
            fc = m.l2_project(prev, w)
            res = float(np.linalg.norm(m.picard_map(P, x, prev, w, m.entropy(fc)) - x))
            tol = max(1e-12 * np.linalg.norm(x), 1e-10)
            print(f"step {step} {k}: {st['iters']} maps, r0 {st['r0']:.3e}, exit {st['exit_reason']}, "
                  f"oracle residual {res:.2e} (tol {tol:.2e})", flush=True)
            assert st["exit_reason"] == 0 and res <= 20 * tol
            r0[k].append(st["r0"])
            prev = x
        last[k] = prev
    assert all(a < b for a, b in zip(r0["mlp"], r0["euler"]))
    print(f"euler vs warm-started state after 2 steps: max|dv| {np.max(np.abs(last['euler'] - last['mlp'])):.2e}")
    # recording and training at N = 65536
    bp = p.bp1
    ws2 = collop.Workspace(bp, bp, ctx=gpu_ctx)
    ml = collop.WarmStart(ws2)
    ml.dataset_alloc(2 * 65536)
    vr, wr = collop.markers("bimodal", 65536, 100, sigma1=0.5 ** 0.5, sigma2=0.5 ** 0.5, peak=2.0)
    prm = p.step_params()
    prm.record_training = 1
    collop.Simulation(ws2, vr, wr).run(prm, 2)
    assert ml.dataset_size() == 2 * 65536
    ml.set_params(W.glorot_init(0))
    ml.fit_norm()
    losses = ml.train(epochs=3, batch=65536, lr=1e-3, seed=1)
    print("training loss per epoch", losses)
    assert all(np.isfinite(losses)) and losses[-1] < losses[0]
