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
        self.w = w
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

    def diagnostics(self) -> dict:
        """The run record's fp_minus_fs and neg_part of the resident state
        (main.jl:306-307, functions.jl:300-343)."""
        self._own()
        out = np.zeros(2)
        self.ws.ctx.check(self.ws.ctx.lib.collop_state_diagnostics(self.ws.ctx.h, _lib.dptr(out)))
        return {"fp_minus_fs": float(out[0]), "neg_part": float(out[1])}


def profile(ws: Workspace, enable: bool):
    """Toggle HIP-event timing of the pair kernels; returns (ms_total[4], launches[4])
    accumulated since the previous call (index 0 FP64, 1 FP32c, 2 FP64 + metric,
    3 FP32c + metric)."""
    ms = np.zeros(4)
    n = (C.c_long * 4)()
    ws.ctx.check(ws.ctx.lib.collop_profile(ws.ctx.h, int(enable), _lib.dptr(ms), n))
    return ms, [n[i] for i in range(4)]


def set_pair_tile(ws: Workspace, gamma_per_lane: int = 1, slices: int = 0, newton: int = 1):
    """Pin the ordered pair-sum tile (slices = 0: automatic) and select the
    ordered sum; see collop_set_pair_tile."""
    ws.ctx.check(ws.ctx.lib.collop_set_pair_tile(ws.ctx.h, gamma_per_lane, slices, newton))


def set_pair_mode(ws: Workspace, symmetric: bool = True, tiles: int = 0):
    """Ordered pair sum (every ordered pair, as the reference's GPU kernel) or
    the symmetric one (every unordered pair once, the default); tiles = waves
    per block of the symmetric sum (0: automatic).  See collop_set_pair_mode."""
    mode = _lib.PAIR_SYMMETRIC if symmetric else _lib.PAIR_ORDERED
    ws.ctx.check(ws.ctx.lib.collop_set_pair_mode(ws.ctx.h, mode, tiles))


def set_fused_maps(ws: Workspace, on: bool = True):
    """Fused small kernels (the default) or one launch per stage in the
    frozen-metric inner maps and the Lenard-Bernstein Anderson maps; both give
    the same bits.  See collop_set_fused_maps."""
    ws.ctx.check(ws.ctx.lib.collop_set_fused_maps(ws.ctx.h, 1 if on else 0))


def set_weight_path(ws: Workspace, on: bool = True):
    """The equal-weight form of the symmetric pair sums (on, the default) or
    the general form: the same bits.  See collop_set_weight_path."""
    ws.ctx.check(ws.ctx.lib.collop_set_weight_path(ws.ctx.h, 1 if on else 0))


def set_device_solve(ws: Workspace, on: bool = True):
    """Anderson control of Landau steps on the device (the default: the host
    only launches maps) or on the host (every map's statistics read back);
    both give the same bits.  See collop_set_device_solve."""
    ws.ctx.check(ws.ctx.lib.collop_set_device_solve(ws.ctx.h, 1 if on else 0))


def pair_mode(ws: Workspace, N: int):
    """(symmetric, tiles, slots) a pair sum over N markers runs with now."""
    mode, tiles, slots = C.c_int(), C.c_int(), C.c_int()
    ws.ctx.check(ws.ctx.lib.collop_get_pair_mode(ws.ctx.h, int(N), C.byref(mode), C.byref(tiles),
                                                 C.byref(slots)))
    return mode.value == _lib.PAIR_SYMMETRIC, tiles.value, slots.value


class StreamEnsemble:
    """Independent runs on one mesh advanced by the single-run device step,
    several at once: each of ``workers`` contexts owns its own HIP stream and
    buffers, and a host thread steps that context's share of the runs (dealt
    largest N² first to the least-loaded worker), so the small O(N) kernels and
    the pair sums of different runs overlap on the GPU.  Unlike :class:`Ensemble`
    (one persistent launch per Anderson iteration) every solver of the
    single-run step is available, defect correction included, and each run's
    result is bitwise that of a :class:`Simulation` of its markers alone."""

    def __init__(self, bp1, bp2, runs, workers: int = 4, device: int = 0, warm_start=None):
        if workers < 1:
            raise ValueError("workers must be >= 1")
        self.v, self.w = [], []
        for v, w in runs:
            v = np.asfortranarray(np.asarray(v, dtype=np.float64))
            w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
            if v.ndim != 2 or v.shape[1] != 2 or w.shape != (v.shape[0],):
                raise ValueError("each run needs v (N, 2) and w (N,)")
            self.v.append(v)
            self.w.append(w)
        self.workers = min(workers, max(1, len(self.v)))
        self.ws = [Workspace(bp1, bp2, device=device) for _ in range(self.workers)]
        if warm_start is not None:
            for ws in self.ws:
                warm_start(ws)
        load = [0.0] * self.workers
        self.share: list[list[int]] = [[] for _ in range(self.workers)]
        for r in sorted(range(len(self.v)), key=lambda r: -float(self.v[r].shape[0]) ** 2):
            k = int(np.argmin(load))
            self.share[k].append(r)
            load[k] += float(self.v[r].shape[0]) ** 2
        # odd workers take their runs smallest first, so that large pair sums and
        # the launch-bound small runs overlap in time (282 against 273 run-steps/s
        # on the 48-run ensemble, three samples each)
        self.share = [s if k % 2 == 0 else s[::-1] for k, s in enumerate(self.share)]

    @property
    def n_runs(self) -> int:
        return len(self.v)

    def step(self, params: _lib.StepParams, n_steps: int = 1) -> list[list[dict]]:
        """n_steps steps of every run; returns stats[step][run]."""
        from concurrent.futures import ThreadPoolExecutor

        stats: list[list[dict]] = [[{} for _ in self.v] for _ in range(n_steps)]

        def work(k: int):
            for r in self.share[k]:
                sim = Simulation(self.ws[k], self.v[r], self.w[r])
                st = sim.run(params, n_steps)
                for s in range(n_steps):
                    stats[s][r] = st[s]
                self.v[r] = sim.markers()

        with ThreadPoolExecutor(self.workers) as ex:
            for f in [ex.submit(work, k) for k in range(self.workers)]:
                f.result()
        return stats

    def markers(self) -> list[np.ndarray]:
        return [v.copy(order="F") for v in self.v]

    def close(self):
        for ws in self.ws:
            ws.ctx.close()


class Ensemble:
    """Independent runs with uneven N on one workspace, advanced together
    (collop_ensemble_*).  Each run solves exactly as a :class:`Simulation` of its
    own markers would with ``set_pair_tile(ws, 1, Ensemble.slices(N), 1)``."""

    def __init__(self, ws: Workspace, runs):
        self.ws = ws
        vs, ws_ = [], []
        for v, w in runs:
            v = np.asfortranarray(np.asarray(v, dtype=np.float64))
            w = np.ascontiguousarray(np.asarray(w, dtype=np.float64))
            if v.ndim != 2 or v.shape[1] != 2 or w.shape != (v.shape[0],):
                raise ValueError("each run needs v (N, 2) and w (N,)")
            vs.append(v)
            ws_.append(w)
        self.n = np.array([v.shape[0] for v in vs], dtype=np.int64)
        vcat = np.concatenate([v.ravel(order="F") for v in vs]) if vs else np.zeros(0)
        wcat = np.concatenate(ws_) if ws_ else np.zeros(0)
        ws.activate()
        ctx = ws.ctx
        ctx.check(ctx.lib.collop_ensemble_set(ctx.h, len(vs), self.n.ctypes.data_as(C.POINTER(C.c_int64)),
                                              _lib.dptr(vcat), _lib.dptr(wcat)))
        ctx.ens_owner = self
        self._mesh = ws._token

    @property
    def n_runs(self) -> int:
        return len(self.n)

    @staticmethod
    def slices(n: int) -> int:
        return int(_lib.load().collop_ensemble_slices(int(n)))

    def _own(self):
        ctx = self.ws.ctx
        if getattr(ctx, "ens_owner", None) is not self or getattr(ctx, "mesh_owner", None) is not self._mesh:
            raise _lib.CollopError(_lib.ERR_NO_STATE, "this ensemble's device state was replaced (another ensemble "
                                   "or mesh on the same context)")

    def step(self, params: _lib.StepParams, n_steps: int = 1) -> list[list[dict]]:
        """n_steps steps of every run; returns stats[step][run]."""
        self._own()
        R = self.n_runs
        stats = (_lib.StepStats * max(n_steps * R, 1))()
        self.ws.ctx.check(self.ws.ctx.lib.collop_ensemble_step(self.ws.ctx.h, C.byref(params), n_steps, stats))
        return [[stats[s * R + r].as_dict() for r in range(R)] for s in range(n_steps)]

    def markers(self) -> list[np.ndarray]:
        self._own()
        flat = np.empty(2 * int(self.n.sum()))
        self.ws.ctx.check(self.ws.ctx.lib.collop_ensemble_get(self.ws.ctx.h, _lib.dptr(flat)))
        out, o = [], 0
        for n in self.n:
            out.append(flat[o:o + 2 * n].reshape((n, 2), order="F").copy(order="F"))
            o += 2 * n
        return out

    def coeffs(self) -> np.ndarray:
        self._own()
        fc = np.empty((self.n_runs, self.ws.n_dofs))
        self.ws.ctx.check(self.ws.ctx.lib.collop_ensemble_coeffs(self.ws.ctx.h, _lib.dptr(fc)))
        return fc

    def set_tile(self, gamma_tile: int = 256, waves_per_cu: int = 8, use_graph: bool = True):
        self._own()
        self.ws.ctx.check(self.ws.ctx.lib.collop_ensemble_set_tile(self.ws.ctx.h, gamma_tile, waves_per_cu,
                                                                    int(use_graph)))

    def tune(self):
        """Autotune the persistent pair sum; returns (gamma_tile, waves_per_cu, ms[4][4])."""
        self._own()
        bt, wv = C.c_int(), C.c_int()
        ms = np.zeros(16)
        self.ws.ctx.check(self.ws.ctx.lib.collop_ensemble_tune(self.ws.ctx.h, C.byref(bt), C.byref(wv),
                                                                _lib.dptr(ms)))
        return bt.value, wv.value, ms.reshape(4, 4)
