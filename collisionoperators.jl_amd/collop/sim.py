"""Device-resident time stepping: the loop of main.jl:226-307 (Landau,
solver = :anderson) with every map, reduction and projection on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import Workspace


class Simulation:
    """Markers resident in HBM; ``run`` advances them by implicit steps."""

    def __init__(self, ws: Workspace, v, w):
        self.ws = ws
        v = np.asfortranarray(np.asarray(v, dtype=np.float64))
        w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
        if v.ndim != 2 or v.shape[1] != 2 or w.shape != (v.shape[0],):
            raise ValueError("v must be (N, 2) and w (N,)")
        self.N = v.shape[0]
        ws.activate()
        ws.ctx.check(ws.ctx.lib.collop_state_set(ws.ctx.h, _lib.dptr(v), _lib.dptr(w), self.N))
        ws.ctx.state_owner = self

    def _own(self):
        ctx = self.ws.ctx
        if getattr(ctx, "state_owner", None) is not self or getattr(ctx, "mesh_owner", None) is not self.ws._token:
            raise _lib.CollopError(_lib.ERR_NO_STATE, "this simulation's device state was replaced by another "
                                   "workspace or simulation on the same context")

    def run(self, params: _lib.StepParams, n_steps: int) -> list[dict]:
        self._own()
        stats = (_lib.StepStats * max(n_steps, 1))()
        self.ws.ctx.check(self.ws.ctx.lib.collop_run_steps(self.ws.ctx.h, C.byref(params), n_steps, stats))
        return [stats[i].as_dict() for i in range(n_steps)]

    def markers(self) -> np.ndarray:
        self._own()
        v = np.empty((self.N, 2), order="F")
        self.ws.ctx.check(self.ws.ctx.lib.collop_state_get(self.ws.ctx.h, _lib.dptr(v)))
        return v

    def coeffs(self) -> np.ndarray:
        self._own()
        fc = np.empty(self.ws.n_dofs)
        self.ws.ctx.check(self.ws.ctx.lib.collop_state_coeffs(self.ws.ctx.h, _lib.dptr(fc)))
        return fc


def profile(ws: Workspace, enable: bool):
    """Toggle HIP-event timing of the pair kernel; returns (ms_total[2], launches[2])
    accumulated since the previous call (index 0 FP64, 1 FP32c)."""
    ms = np.zeros(2)
    n = (C.c_long * 2)()
    ws.ctx.check(ws.ctx.lib.collop_profile(ws.ctx.h, int(enable), _lib.dptr(ms), n))
    return ms, [n[0], n[1]]


def set_pair_tile(ws: Workspace, gamma_per_lane: int = 1, slices: int = 0, newton: int = 1):
    """Pin the pair-sum tile (slices = 0: automatic); see collop_set_pair_tile."""
    ws.ctx.check(ws.ctx.lib.collop_set_pair_tile(ws.ctx.h, gamma_per_lane, slices, newton))
