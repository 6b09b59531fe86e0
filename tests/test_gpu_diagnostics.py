"""Run-record diagnostics on the device against the oracle: the negative-part
L1 norm and ||f_s - f_p||_2 (functions.jl:300-343), at the hook level and on
the resident state, and the run record written from device steps (CSV rows,
checkpoint, resume; io.jl, main.jl:296-346).  The quadrature sums run in a
different order on the device, so the bar is 1e-12 relative.
"""
import numpy as np
import pytest

import collop
import oracle as O
from conftest import uniform_bp

pytestmark = pytest.mark.gpu

NONUNI = np.array([-6.0, -5.0, -4.0, -3.0, -1.5, 0.0, 0.7, 2.0, 4.0, 5.5, 6.0])


def _close(a, b, rel=1e-12, floor=1e-15):
    return abs(a - b) <= rel * abs(b) + floor


@pytest.mark.parametrize("mesh", ["uniform32", "nonuniform"])
def test_hook_diagnostics_match_oracle(gpu_ctx, mesh):
    bp1, bp2 = (uniform_bp(-6, 6, 32),) * 2 if mesh == "uniform32" else (NONUNI, uniform_bp(-6, 6, 9))
    m = O.Mesh(bp1, bp2)
    ws = collop.Workspace(bp1, bp2, ctx=gpu_ctx)
    rng = np.random.default_rng(21)
    N = 20000
    v, w = collop.markers("bimodal", N, 5, sigma1=np.sqrt(0.5), sigma2=np.sqrt(0.5), peak=2.0, weight_jitter=0.3)
    v = np.asfortranarray(v)
    v[:40] = rng.uniform(-7.0, 7.0, (40, 2))      # some outside the box
    v[40] = (bp1[0], 0.3)                         # on the boundary: outside the open box
    v[41] = (0.1, bp2[-1])
    fc = m.l2_project(v, w)
    cases = [("projection", fc), ("signed noise", rng.standard_normal(m.n_dofs) * 0.02 + 0.002)]
    for name, c in cases:
        neg_ref, fp_ref = m.negative_part_l1(c), m.fs_minus_fp_l2(c, v, w)
        neg = collop.compute_negative_part_l1_gpu(ws, c)
        fp = collop.compute_fs_minus_fp_l2_gpu(ws, c, v, w)
        print(f"{mesh} {name}: neg {neg:.6e} (oracle {neg_ref:.6e}), fp {fp:.6e} (oracle {fp_ref:.6e})")
        assert _close(neg, neg_ref) and _close(fp, fp_ref)
    assert collop.compute_negative_part_l1_gpu(ws, cases[1][1]) > 1e-3
    # no markers: f_p = 0, so the norm is ||f_s||_2
    fp0 = collop.compute_fs_minus_fp_l2_gpu(ws, fc, np.zeros((0, 2)), np.zeros(0))
    assert _close(fp0, m.fs_minus_fp_l2(fc, np.zeros((0, 2)), np.zeros(0)))


def test_state_diagnostics_and_run_record(gpu_ctx, tmp_path):
    bp = uniform_bp(-6, 6, 16)
    m = O.Mesh(bp, bp)
    v, w = collop.markers("gaussian", 2048, 42)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    sim = collop.Simulation(ws, v, w)
    d0 = sim.diagnostics()
    fc0 = sim.coeffs()
    assert _close(d0["neg_part"], m.negative_part_l1(fc0))
    assert _close(d0["fp_minus_fs"], m.fs_minus_fp_l2(fc0, v, w))

    params = collop.config("gaussian_n4096", DT=0.01).step_params()
    rec = collop.RunRecord("gpu", str(tmp_path))
    (p1, p2), E0 = O.momentum_energy(v, w)
    rec.write_initial(m.entropy(fc0), E0, (p1, p2))
    out = collop.run_recorded(sim, params, 3, rec, checkpoint_every=2, checkpoint_dir=str(tmp_path))
    assert [o["exit_reason"] for o in out] == [0, 0, 0]
    vg, fcg = sim.markers(), sim.coeffs()
    assert _close(out[-1]["neg_part"], m.negative_part_l1(fcg))
    assert _close(out[-1]["fp_minus_fs"], m.fs_minus_fp_l2(fcg, vg, w))

    lines = (tmp_path / "conservation_history_gpu.csv").read_text().splitlines()
    assert lines[0].split(",") == collop.CONS_COLS and len(lines) == 5
    last = dict(zip(collop.CONS_COLS, lines[-1].split(",")))
    assert int(last["step"]) == 3 and abs(float(last["time"]) - 0.03) < 1e-15
    assert float(last["fp_minus_fs"]) == out[-1]["fp_minus_fs"]
    assert float(last["entropy"]) == out[-1]["entropy"] and float(last["dt"]) == 0.01
    assert abs(float(last["energy"]) - E0) <= 1e-12 * E0

    # the last checkpoint holds the resident state; a resumed run re-projects it bitwise
    ck = collop.load_checkpoint("gpu", "auto", str(tmp_path))
    assert ck["step"] == 3 and len(ck["entropy_history"]) == 3
    np.testing.assert_array_equal(ck["v_particles"], vg)
    np.testing.assert_array_equal(ck["f_coeffs"], fcg)
    sim2 = collop.Simulation(ws, ck["v_particles"], ck["w_particles"])
    np.testing.assert_array_equal(sim2.coeffs(), ck["f_coeffs"])
    assert sim2.diagnostics() == {"fp_minus_fs": out[-1]["fp_minus_fs"], "neg_part": out[-1]["neg_part"]}
