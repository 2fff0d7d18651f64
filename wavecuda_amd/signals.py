"""Synthetic test signals: Donoho-Johnstone doppler plus Gaussian noise.

doppler: s(t) = sqrt(t(1-t)) sin(2 pi 1.05 / (t + 0.05)), t = i/n, i = 1..n
noise:   i.i.d. N(0, (sd(s)/snr)^2), sd the sample standard deviation, drawn
         from numpy's PCG64 generator with a fixed seed.
"""
from __future__ import annotations

import numpy as np


def doppler(n: int) -> np.ndarray:
    t = np.arange(1, n + 1, dtype=np.float64) / n
    return np.sqrt(t * (1.0 - t)) * np.sin(2.0 * np.pi * 1.05 / (t + 0.05))


def noisy_doppler(n: int, snr: float = 5.0, seed: int = 10):
    """Returns (noisy, clean) float64 arrays of length n."""
    clean = doppler(n)
    sd = clean.std(ddof=1) / snr
    rng = np.random.Generator(np.random.PCG64(seed))
    return clean + sd * rng.standard_normal(n), clean


def white_noise(n: int, seed: int = 1) -> np.ndarray:
    return np.random.Generator(np.random.PCG64(seed)).standard_normal(n)


def batch_workload(count: int = 8192, seed: int = 2024, kmin: int = 8, kmax: int = 16):
    """Mixed batch of noisy doppler signals: per signal k ~ U{kmin..kmax}
    (n = 2^k), filter ~ U{Haar, D4, C6, LA8}, transform ~ U{DWT, MODWT},
    SNR ~ U[1, 10] and its own noise stream.  The DWT runs at full depth
    (levels 0), the MODWT at J = min(8, k - 2).

    Returns (x, lens, filters, ttypes, levels): x the float64 signals packed
    back to back, the rest int32 arrays of one entry per signal."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = rng.integers(kmin, kmax + 1, size=count)
    filters = rng.integers(1, 5, size=count).astype(np.int32)
    ttypes = rng.integers(0, 2, size=count).astype(np.int32)
    snr = rng.uniform(1.0, 10.0, size=count)
    lens = (1 << k).astype(np.int32)
    levels = np.where(ttypes == 1, np.minimum(8, k - 2), 0).astype(np.int32)
    x = np.empty(int(lens.sum()), dtype=np.float64)
    off = 0
    cache = {}
    for i in range(count):
        n = int(lens[i])
        if n not in cache:
            c = doppler(n)
            cache[n] = (c, c.std(ddof=1))
        clean, sd = cache[n]
        x[off:off + n] = clean + (sd / snr[i]) * rng.standard_normal(n)
        off += n
    return x, lens, filters, ttypes, levels
