"""Learned warm start: the per-marker MLP of the nn_warmstart_correct! hook site
(main.jl:164-166, 252-254; Parameters.jl:108-117), trained on the GPU by
libcollop.so's own forward/backward kernels on trajectories the solver generates.

Workflow (everything on the device; host code only orchestrates):
  1. ``dataset_alloc`` and run steps with ``record_training = 1``: every marker of
     every step contributes features (v, G, f_s, |grad f_s|) and the target
     (v_{n+1} - v_n)/dt - vdot_euler;
  2. ``fit_norm`` sets the feature / target normalisation from the data;
  3. ``train`` runs Adam on the batch-mean squared error;
  4. steps with ``warmstart = WARM_MLP`` apply v1 = v + dt (vdot_euler + y * scale).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import Workspace

NPARAM = 4738
SIZES = [("W1", (64, 6)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (2, 64)), ("b3", (2,))]


def glorot_init(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases; float32[4738] in the library's layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-lim, lim, shape).astype(np.float32).reshape(-1))
        else:
            parts.append(np.zeros(int(np.prod(shape)), np.float32))
    return np.concatenate(parts)


class WarmStart:
    """The MLP state held by one workspace's device context."""

    def __init__(self, ws: Workspace):
        self.ws = ws
        self.norm = None

    @property
    def _c(self):
        return self.ws.ctx

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._c.check(self._c.lib.collop_mlp_set_params(self._c.h, _lib.fptr(p), len(p)))

    def params(self) -> np.ndarray:
        p = np.zeros(NPARAM, np.float32)
        self._c.check(self._c.lib.collop_mlp_get_params(self._c.h, _lib.fptr(p), NPARAM))
        return p

    def set_norm(self, mu, sd, out_scale):
        mu = np.ascontiguousarray(mu, dtype=np.float32)
        sd = np.ascontiguousarray(sd, dtype=np.float32)
        os_ = np.ascontiguousarray(out_scale, dtype=np.float32)
        self._c.check(self._c.lib.collop_mlp_set_norm(self._c.h, _lib.fptr(mu), _lib.fptr(sd), _lib.fptr(os_)))
        self.norm = (mu, sd, os_)

    def dataset_alloc(self, capacity: int):
        self._c.check(self._c.lib.collop_mlp_dataset_alloc(self._c.h, int(capacity)))

    def dataset_size(self) -> int:
        n = C.c_int64(0)
        self._c.check(self._c.lib.collop_mlp_dataset_size(self._c.h, C.byref(n)))
        return n.value

    def dataset_read(self, first: int, count: int) -> np.ndarray:
        out = np.zeros((count, 8), np.float32)
        self._c.check(self._c.lib.collop_mlp_dataset_read(self._c.h, int(first), int(count), _lib.fptr(out)))
        return out

    def fit_norm(self, max_samples: int = 1 << 20):
        """Feature mean / std and target RMS from an evenly strided sample of the dataset."""
        n = self.dataset_size()
        if n == 0:
            raise _lib.CollopError(_lib.ERR_INVALID, "empty dataset")
        chunks = []
        step = max(1, n // max_samples)
        block = 65536
        for first in range(0, n, block * step):
            chunks.append(self.dataset_read(first, min(block, n - first)))
        d = np.concatenate(chunks).astype(np.float64)
        mu = d[:, :6].mean(axis=0)
        sd = d[:, :6].std(axis=0) + 1e-12
        os_ = np.sqrt((d[:, 6:8] ** 2).mean(axis=0)) + 1e-30
        self.set_norm(mu, sd, os_)
        return mu, sd, os_

    def train(self, epochs: int, batch: int = 65536, lr: float = 1e-3, seed: int = 0) -> list[float]:
        loss = np.zeros(max(epochs, 1))
        self._c.check(self._c.lib.collop_mlp_train(self._c.h, epochs, batch, lr, seed, _lib.dptr(loss)))
        return loss[:epochs].tolist()

    def eval(self, rec: np.ndarray, mode: int):
        rec = np.ascontiguousarray(rec, dtype=np.float32)
        n = rec.shape[0]
        y = np.zeros((n, 2), np.float32)
        loss = np.zeros(1)
        grad = np.zeros(NPARAM, np.float32)
        self._c.check(self._c.lib.collop_mlp_eval(self._c.h, _lib.fptr(rec), n, mode, _lib.fptr(y), _lib.dptr(loss),
                                                  _lib.fptr(grad)))
        return y, float(loss[0]), grad

    def save(self, path: str):
        mu, sd, os_ = self.norm
        np.savez(path, params=self.params(), mu=mu, sd=sd, out_scale=os_)

    def load(self, path: str):
        d = np.load(path, allow_pickle=False)
        self.set_params(d["params"])
        self.set_norm(d["mu"], d["sd"], d["out_scale"])


def training_runs(n_runs: int = 8, N: int = 65536, seed0: int = 100):
    """The seeded training runs: the isotropic, bimodal (+-2, cov 0.5 I) and
    bi-Maxwellian (T_x = 2, T_y = 0.5) ICs in turn, equal weights."""
    from .ic import ENSEMBLE_ICS, markers
    out = []
    for r in range(n_runs):
        ic, s1, s2 = ENSEMBLE_ICS[r % 3]
        v, w = markers(ic, N, seed0 + r, sigma1=s1, sigma2=s2, peak=2.0)
        out.append((ic, v, w))
    return out


def record(ws: Workspace, runs, steps: int, params: _lib.StepParams, log=None) -> int:
    """Step every run `steps` times with record_training = 1 (one sample per
    marker per step appended to the device dataset); returns the sample count."""
    from .sim import Simulation
    ml = WarmStart(ws)
    N = runs[0][1].shape[0]
    ml.dataset_alloc(len(runs) * steps * N)
    prm = _lib.StepParams.from_buffer_copy(params)
    prm.record_training = 1
    prm.warmstart = _lib.WARM_EULER
    for i, (ic, v, w) in enumerate(runs):
        sim = Simulation(ws, v, w)
        st = sim.run(prm, steps)
        if log:
            log(f"  run {i} ({ic}): {steps} steps, mean maps {np.mean([s['iters'] for s in st]):.1f}")
    return ml.dataset_size()


def evaluate(ws: Workspace, v, w, params: _lib.StepParams, steps: int) -> dict:
    """`steps` steps from (v, w): per-step maps, initial residual r0, wall time."""
    import time

    from .sim import Simulation
    sim = Simulation(ws, v, w)
    t0 = time.perf_counter()
    st = sim.run(params, steps)
    el = time.perf_counter() - t0
    return {"steps_per_s": steps / el, "mean_maps": float(np.mean([s["iters"] for s in st])),
            "mean_inner": float(np.mean([s["inner_maps"] for s in st])),
            "mean_r0": float(np.mean([s["r0"] for s in st])),
            "median_r0": float(np.median([s["r0"] for s in st])),
            "all_converged": all(s["exit_reason"] == 0 for s in st), "markers": sim.markers()}


MODEL_DIR = __import__("os").path.join(__import__("os").path.dirname(__file__), "models")
# trained by `bench.py --report warmstart` (8 runs x 200 recorded steps at N = 65536 on
# the 32 x 32 mesh over [-6, 6]^2, dt = 0.02; 20 Adam epochs, lr 1e-3)
PACKAGED_MODEL = __import__("os").path.join(MODEL_DIR, "warmstart_mlp_32x32.npz")


def load_packaged(ws: Workspace, path: str | None = None) -> "WarmStart":
    """Load the packaged warm-start model onto ws's context."""
    ml = WarmStart(ws)
    ml.load(path or PACKAGED_MODEL)
    return ml
