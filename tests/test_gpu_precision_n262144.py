"""The precision form at its full size: the anisotropic bi-Maxwellian with
N = 262144 markers (T_x = 2, T_y = 0.5, weights proportional to 1 + 0.2u) on
the 64x64 mesh over [-8, 8]^2.  At this N the symmetric pair sum runs with 8
waves per block and all 256 partial slots.

Parity against the oracle on a subsample of 512 targets (every source), for
the FP64 and the FP32-compensated pair sums; the projection coefficients in
full; and, over all 262144 markers, the size-independent invariants of the
pair sum: sum_g w_g vdot_g = 0 (each pair's term is applied with opposite
signs) and sum_g w_g v_g . vdot_g = 0 (U(d) d = 0), both to roundoff.
"""
import numpy as np
import pytest

import collop
import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu

N = 262144
ROWS = 512


@pytest.fixture(scope="module")
def big(gpu_ctx):
    p = collop.config("bimaxwellian_n262144")
    v, w = collop.initial_markers(p)
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    G = m.compute_G(v, m.mass_solve(m.compute_r(fc)))
    ws = collop.Workspace(p.bp1, p.bp2, ctx=gpu_ctx)
    ref = O.collision_rows(m.box, v, w, G, 0, ROWS)
    return dict(v=v, w=w, m=m, fc=fc, G=G, ws=ws, ref=ref[:ROWS])


def _invariants(v, w, dot):
    """Relative residuals of sum w vdot and sum w v.vdot, against the sums of
    the magnitudes of their terms."""
    p = (w[:, None] * dot).sum(axis=0)
    p_scale = (np.abs(w[:, None] * dot)).sum(axis=0)
    e = float((w * (v * dot).sum(axis=1)).sum())
    e_scale = float(np.abs(w[:, None] * v * dot).sum())
    return float(np.max(np.abs(p) / p_scale)), abs(e) / e_scale


def test_pair_mode_at_n262144(big):
    assert collop.pair_mode(big["ws"], N) == (True, 8, 256)


def test_collision_f64_n262144(big):
    out = np.zeros((N, 2), order="F")
    collop.compute_collision_gpu(big["ws"], out, big["v"], big["w"], big["G"])
    e = relerr(out[:ROWS], big["ref"])
    dp, de = _invariants(big["v"], big["w"], out)
    print(f"N=262144 FP64: relerr ({ROWS} targets) {e:.2e}; invariants {dp:.1e} {de:.1e}")
    assert e <= 1e-12
    assert dp <= 1e-15 and de <= 1e-15  # measured 2.2e-18 and 8.5e-19


def test_collision_f32c_n262144(big):
    out = np.zeros((N, 2), order="F")
    collop.compute_collision_gpu32(big["ws"], out, big["v"], big["w"], big["G"])
    e = relerr(out[:ROWS], big["ref"])
    dp, de = _invariants(big["v"], big["w"], out)
    print(f"N=262144 FP32c: relerr ({ROWS} targets) {e:.2e}; invariants {dp:.1e} {de:.1e}")
    assert e <= 2e-6
    # each pair term is formed once and applied to both markers, and the sums
    # are compensated: both invariants hold far below FP32 rounding of a term
    assert dp <= 1e-8 and de <= 1e-8  # measured 5.9e-11 and 1.2e-10


def test_projection_n262144(big):
    fc = np.zeros_like(big["fc"])
    collop.l2_project_gpu(big["ws"], fc, big["v"], big["w"])
    e = relerr(fc, big["fc"])
    print(f"N=262144 projection relerr {e:.2e}")
    assert e <= 1e-13
