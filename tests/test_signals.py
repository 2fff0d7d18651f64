"""Synthetic inputs (no device needed): the doppler formula, the noise
level rule, and the 8192-signal batch workload's composition."""
import numpy as np

from wavecuda_amd.signals import batch_workload, doppler, noisy_doppler


def test_doppler_formula():
    n = 1024
    t = np.arange(1, n + 1) / n
    ref = np.sqrt(t * (1 - t)) * np.sin(2 * np.pi * 1.05 / (t + 0.05))
    np.testing.assert_array_equal(doppler(n), ref)
    assert doppler(n)[-1] == 0.0  # t = 1


def test_noise_level_and_seed():
    x, clean = noisy_doppler(4096, 5.0, seed=10)
    x2, _ = noisy_doppler(4096, 5.0, seed=10)
    np.testing.assert_array_equal(x, x2)
    sd = np.std(x - clean, ddof=1)
    assert abs(sd / (clean.std(ddof=1) / 5.0) - 1.0) < 0.05


def test_batch_workload_composition():
    x, lens, filters, ttypes, levels = batch_workload(count=512, seed=7)
    assert len(lens) == 512 and x.size == int(lens.astype(np.int64).sum())
    k = np.log2(lens).astype(int)
    assert np.all((k >= 8) & (k <= 16)) and np.all(lens == (1 << k))
    assert set(np.unique(filters)) <= {1, 2, 3, 4} and set(np.unique(ttypes)) <= {0, 1}
    # DWT at full depth (levels 0), MODWT at J = min(8, k - 2)
    assert np.all(levels[ttypes == 0] == 0)
    assert np.all(levels[ttypes == 1] == np.minimum(8, k[ttypes == 1] - 2))
    xb, lb, fb, tb, vb = batch_workload(count=512, seed=7)
    np.testing.assert_array_equal(xb, x)


def test_full_workload_size():
    _, lens, _, _, _ = batch_workload()
    assert len(lens) == 8192
    assert 110e6 < lens.astype(np.int64).sum() < 130e6
