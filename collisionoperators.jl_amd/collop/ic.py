"""Seeded initial markers (the IC block of main.jl:91-103).

The reference draws with Julia's RNG; this library uses numpy's PCG64 with
the given seed.  Markers are float64 and column-major (N x 2, v1 then v2), so
the oracle and the device consume identical bytes.
"""
from __future__ import annotations

import numpy as np


def initial_markers(p) -> tuple[np.ndarray, np.ndarray]:
    """Return (v, w) for a SimParameters-like ``p`` (fields N_PARTICLES, ic,
    σ1, σ2, v1_peak, weight_jitter, seed)."""
    return markers(p.ic, p.N_PARTICLES, p.seed, sigma1=p.σ1, sigma2=p.σ2, peak=p.v1_peak,
                   weight_jitter=p.weight_jitter)


def markers(ic: str, N: int, seed: int, sigma1: float = 1.0, sigma2: float = 1.0, peak: float = 2.0,
            weight_jitter: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.Generator(np.random.PCG64(seed))
    v = np.empty((N, 2), order="F")
    if ic == "bimodal":
        # balanced halves at -peak / +peak in v1 (main.jl:96-98)
        half = N // 2
        centers = np.concatenate([np.full(half, -peak), np.full(N - half, peak)])
        v[:, 0] = centers + sigma1 * rng.standard_normal(N)
    elif ic in ("gaussian", "bimaxwellian"):
        v[:, 0] = sigma1 * rng.standard_normal(N)
    else:
        raise ValueError(f"unknown ic {ic!r}")
    v[:, 1] = sigma2 * rng.standard_normal(N)
    if weight_jitter > 0.0:
        u = rng.uniform(-1.0, 1.0, N)
        w = 1.0 + weight_jitter * u
        w /= w.sum()
    else:
        w = np.full(N, 1.0 / N)
    return v, w
