"""The multi-stream ensemble (several contexts, each with its own HIP stream,
stepped by host threads): every run's result is bitwise that of the same run
stepped alone, with the Anderson path's solver (FP32c then FP64 maps)."""
import numpy as np
import pytest

import collop
from conftest import uniform_bp

pytestmark = pytest.mark.gpu


def test_stream_ensemble_matches_single_runs(gpu_ctx):
    bp = uniform_bp(-6, 6, 32)
    ics = ["gaussian", "bimodal", "bimaxwellian", "gaussian", "bimodal"]
    sizes = [700, 1500, 2300, 3100, 900]
    runs = []
    for i, (ic, n) in enumerate(zip(ics, sizes)):
        sig = {"gaussian": (1.0, 1.0), "bimodal": (np.sqrt(0.5), np.sqrt(0.5)),
               "bimaxwellian": (np.sqrt(2.0), np.sqrt(0.5))}[ic]
        runs.append(collop.markers(ic, n, 10 + i, sigma1=sig[0], sigma2=sig[1], peak=2.0))
    params = collop.anderson_config("bimodal_n65536", DT=0.01).step_params(precision=collop.PREC_F32C)
    ens = collop.StreamEnsemble(bp, bp, runs, workers=2)
    try:
        stats = ens.step(params, 2)
        got = ens.markers()
    finally:
        ens.close()
    assert sorted(r for share in ens.share for r in share) == list(range(len(runs)))
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    for r, (v, w) in enumerate(runs):
        sim = collop.Simulation(ws, v, w)
        st = sim.run(params, 2)
        np.testing.assert_array_equal(got[r], sim.markers())
        assert [s["iters"] for s in st] == [stats[k][r]["iters"] for k in range(2)]
        assert all(stats[k][r]["exit_reason"] == 0 for k in range(2))
