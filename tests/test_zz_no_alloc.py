"""The allocation-free pair sum (collop_collision_f64_noalloc) against the
allocating entry point collop_collision_f64 on the same markers, through the
C-ABI with caller-owned device arrays: two calls in a row, the second with a
smaller N (the shared scratch keeps the first call's contents), both within
the pair sum's FP64 bar 1e-12 (normwise); and N beyond the scratch's capacity
is rejected with a message."""
import ctypes as C

import numpy as np
import pytest

import collop
import oracle as O
from conftest import relerr, uniform_bp

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class DevArray:
    def __init__(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
        self.n = a.size
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(8 * self.n)) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0  # H2D

    def get(self):
        out = np.zeros(self.n)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n), 2) == 0  # D2H
        return out

    def free(self):
        hip().hipFree(self.p)


def test_noalloc_pair_sum_matches_the_allocating_entry_point(gpu_ctx):
    bp = uniform_bp(-6, 6, 32)
    m = O.Mesh(bp, bp)
    ws = collop.Workspace(bp, bp, ctx=gpu_ctx)
    for N, seed in ((65536, 1), (3001, 2)):
        v, w = collop.markers("bimodal", N, seed, sigma1=0.5 ** 0.5, sigma2=0.5 ** 0.5, peak=2.0)
        G = m.compute_G(v, m.mass_solve(m.compute_r(m.l2_project(v, w))))
        ref = np.zeros((N, 2), order="F")
        collop.compute_collision_gpu(ws, ref, v, w, G)
        dv, dw, dG = DevArray(v), DevArray(w), DevArray(G)
        dout = DevArray(np.full((N, 2), np.nan))
        print(f"N={N}: allocation-free pair sum", flush=True)
        rc = gpu_ctx.lib.collop_collision_f64_noalloc(gpu_ctx.h, dv.p, dw.p, dG.p, dout.p, N, None)
        assert rc == 0, gpu_ctx.lib.collop_last_error(gpu_ctx.h)
        assert hip().hipDeviceSynchronize() == 0
        got = dout.get().reshape(2, N).T
        err = relerr(got, ref)
        print(f"N={N}: max |noalloc - collop_collision_f64| / max |ref| = {err:.2e}")
        assert err <= 1e-12
        for a in (dv, dw, dG, dout):
            a.free()
    # beyond the scratch: an error with a message, nothing launched
    rc = gpu_ctx.lib.collop_collision_f64_noalloc(gpu_ctx.h, None, None, None, None, 65537, None)
    assert rc == collop._lib.ERR_INVALID
    assert b"65536" in gpu_ctx.lib.collop_last_error(gpu_ctx.h)
