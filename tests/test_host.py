"""Host logic of the collop package that needs no GPU: the Workspace tables
(checked against the oracle's Borden extraction), parameter overrides
(Parameters.jl:157-192 semantics), the seeded initial conditions and the
named workloads."""
import math

import numpy as np
import pytest

import collop
import oracle as O
from conftest import uniform_bp

NONUNI = np.array([-6.0, -5.0, -4.0, -3.0, -1.5, 0.0, 0.7, 2.0, 4.0, 5.5, 6.0])


@pytest.mark.parametrize("bp", [uniform_bp(-6, 6, 32), uniform_bp(-8, 8, 64), NONUNI, uniform_bp(0, 1, 1),
                                uniform_bp(0, 1, 2)])
def test_extraction_matches_oracle(bp):
    C, bs, nd = collop.extraction_p2c1(bp)
    Co, bso, ndo = O.extraction_1d(bp, 2, 1)
    assert nd == ndo
    np.testing.assert_array_equal(bs - 1, bso)
    np.testing.assert_allclose(C, Co, rtol=0, atol=4e-16)
    # flat layout of projection_gpu.jl:191-197: C[i + 3 j + 9 e]
    flat = collop.mesh.flat_tables(C)
    e, i, j = 1 if len(bp) > 2 else 0, 2, 1
    assert flat[i + 3 * j + 9 * e] == C[e, i, j]


def test_extraction_rejects_bad_breakpoints():
    with pytest.raises(ValueError):
        collop.extraction_p2c1([0.0, 1.0, 1.0])
    with pytest.raises(ValueError):
        collop.extraction_p2c1([0.0])


def test_workspace_rejects_unsupported_space():
    with pytest.raises(collop.CollopError) as e:
        collop.Workspace(uniform_bp(0, 1, 4), uniform_bp(0, 1, 4), p_deg=3)
    assert e.value.status == collop._lib.ERR_UNSUPPORTED
    with pytest.raises(collop.CollopError):
        collop.Workspace(uniform_bp(0, 1, 4), uniform_bp(0, 1, 4), k_reg=0)


def test_parse_overrides():
    base = collop.SimParameters()
    p = collop.parse_overrides(base, ["--DT=0.004", "--N_STEPS=300", "--use_gonzalez=false", "--warmstart=:nn"])
    assert p.DT == 0.004 and p.N_STEPS == 300 and p.use_gonzalez is False and p.warmstart == "nn"
    assert p.N_PARTICLES == base.N_PARTICLES
    assert collop.parse_overrides(base, []) == base
    with pytest.raises(ValueError):
        collop.parse_overrides(base, ["--no_such_field=1"])
    with pytest.raises(ValueError):
        collop.parse_overrides(base, ["--bp1=1"])
    with pytest.raises(ValueError):
        collop.parse_overrides(base, ["--use_anderson=maybe"])


def test_initial_conditions():
    v, w = collop.markers("bimodal", 1001, 7, sigma1=math.sqrt(0.5), sigma2=math.sqrt(0.5), peak=2.0)
    assert v.flags.f_contiguous and v.shape == (1001, 2)
    assert np.all(w == 1 / 1001)
    assert abs(v[:500, 0].mean() + 2.0) < 0.15 and abs(v[500:, 0].mean() - 2.0) < 0.15
    v2, _ = collop.markers("bimodal", 1001, 7, sigma1=math.sqrt(0.5), sigma2=math.sqrt(0.5), peak=2.0)
    np.testing.assert_array_equal(v, v2)
    _, w = collop.markers("bimaxwellian", 5000, 3, sigma1=math.sqrt(2), sigma2=math.sqrt(0.5), weight_jitter=0.2)
    assert abs(w.sum() - 1.0) < 1e-14 and w.min() >= 0.8 / 5000 / 1.01 and w.max() <= 1.2 / 5000 * 1.01
    with pytest.raises(ValueError):
        collop.markers("nope", 10, 1)


def test_named_workloads():
    c2 = collop.config("bimodal_n65536")
    assert c2.N_PARTICLES == 65536 and c2.DT == 0.02 and len(c2.bp1) == 33 and c2.m_anderson == 5
    assert c2.tol == 1e-12 and c2.abs_floor == 1e-10 and c2.stag_window == 30
    sp = c2.step_params()
    assert sp.precision == collop.PREC_F64 and sp.m_anderson == 5 and math.isinf(sp.restart_factor)
    c3 = collop.config("bimaxwellian_n262144")
    assert c3.N_PARTICLES == 262144 and c3.bp1[0] == -8.0 and len(c3.bp1) == 65
    assert collop.config("gaussian_n4096").DT == 0.05
