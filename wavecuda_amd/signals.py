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
