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


# (ic, sigma1, sigma2) of the three initial conditions an ensemble draws from
ENSEMBLE_ICS = (("gaussian", 1.0, 1.0), ("bimodal", 0.5 ** 0.5, 0.5 ** 0.5), ("bimaxwellian", 2.0 ** 0.5, 0.5 ** 0.5))


def ensemble_markers(n_runs: int = 48, seed: int = 7, n_min: int = 2048, n_max: int = 65536):
    """Seeded ensemble of runs: N log-uniform in [n_min, n_max], the IC drawn
    from isotropic N(0, I), bimodal (+-2, cov 0.5 I) and bi-Maxwellian
    (T_x = 2, T_y = 0.5), equal weights.  Returns [(v, w, ic, N)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for r in range(n_runs):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        ic, s1, s2 = ENSEMBLE_ICS[int(rng.integers(0, 3))]
        v, w = markers(ic, n, seed * 1000 + r, sigma1=s1, sigma2=s2, peak=2.0)
        out.append((v, w, ic, n))
    return out
