"""GPU parity of thresholding, the smoothing statistics and the smoothing
pipeline (manual / universal / cross-validated) against the CPU oracle,
plus the reference's .C entry points called the way R's .C calls them.

Bars: thresholding, medians, MADs and universal thresholds of the same
coefficients bit-exact (order statistics and the reference's own final
operations); the CV interpolation error within 1e-13 (summation order
differs); smoothed outputs within 1e-12 norm-wise in FP64.  Chosen
thresholds agree to 1e-12 relative: the GPU transforms contract products
into FMAs, so their coefficients -- and the MAD taken from them -- can
differ from the CPU's in the last bit (the Haar filter, which has nothing to
contract, matches exactly).
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import _lib as W
from wavecuda_amd import device as D
from wavecuda_amd.signals import noisy_doppler, white_noise

pytestmark = pytest.mark.gpu
FILTERS = [1, 2, 3, 4]


def same_t(a, b):
    return abs(a - b) <= 1e-12 * abs(b)


def _dev(a, dtype=torch.float64):
    return torch.tensor(a, dtype=dtype, device="cuda")


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("ttype", [0, 1])
@pytest.mark.parametrize("hard", [0, 1])
def test_threshold_inplace(filt, ttype, hard):
    n = 4096
    J = O.levels(n, 8 if ttype else 0, filt)
    x, _ = noisy_doppler(n, 3.0, seed=filt)
    c = O.dwt(x, J, filt) if ttype == 0 else O.modwt_fwd(x, J, filt)
    t = float(np.quantile(np.abs(c), 0.6))
    c[3] = t
    c[9] = -t
    for mn, mx in ((0, J - 1), (1, 3), (J - 1, J - 1)):
        ref = O.threshold(c, n, J, ttype, filt, t, hard, mn, mx)
        g = _dev(c)
        D.threshold_(g, n, J, ttype, filt, t, hard, mn, mx)
        np.testing.assert_array_equal(g.cpu().numpy(), ref)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 1000, 1001, 4096, 1 << 18, (1 << 20) + 3])
def test_median_mad_exact(n):
    rng = np.random.Generator(np.random.PCG64(n))
    x = np.round(rng.standard_normal(n) * 64) / 64  # many ties
    g = _dev(x)
    assert D.median(g).item() == O.median(x)
    assert D.mad(g).item() == O.mad(x)


@pytest.mark.parametrize("case", ["exact_ties", "crowded_bucket", "past_2p22"])
def test_median_mad_candidate_overflow(case):
    """More keys under a middle rank's leading digits than the finishing
    workgroup holds: a bucket of one repeated value (resolved from the
    bucket's min = max), a bucket of 1000 distinct values within 2^-30 of
    each other (resolved over the data), and 2^22 + 5 values (three grid
    digits before the compaction)."""
    rng = np.random.Generator(np.random.PCG64(7))
    if case == "exact_ties":
        x = rng.integers(-1, 3, 1 << 20).astype(np.float64)
    elif case == "crowded_bucket":
        x = 1.0 + rng.integers(0, 1000, 1 << 20) * 2.0 ** -40
    else:
        x = np.round(rng.standard_normal((1 << 22) + 5) * 64) / 64
    print(f"median overflow {case} n={x.size}", flush=True)
    g = _dev(x)
    assert D.median(g).item() == O.median(x)
    assert D.mad(g).item() == O.mad(x)
    # FP32: the median is exact; the MAD's deviations round in FP32
    x32 = x.astype(np.float32)
    g32 = _dev(x32, torch.float32)
    assert D.median(g32).item() == O.median(x32.astype(np.float64))
    m = O.mad(x32.astype(np.float64))
    assert abs(D.mad(g32).item() - m) <= 1e-6 * max(m, 1e-30)


def test_median_special_values():
    x = np.array([-0.0, 0.0, -1e-300, 1e300, -np.inf, np.inf, 5e-324, -5e-324, 2.5])
    g = _dev(x)
    assert D.median(g).item() == O.median(x)
    y = np.array([3.0, -2.0, -2.0, 7.0])
    assert D.median(_dev(y)).item() == O.median(y)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 1001, 1 << 18])
def test_median_f32_sizes(n):
    rng = np.random.Generator(np.random.PCG64(n + 5))
    x = (np.round(rng.standard_normal(n) * 64) / 64).astype(np.float32)  # many ties
    g = torch.tensor(x, device="cuda")
    assert D.median(g).item() == O.median(x.astype(np.float64))
    m = O.mad(x.astype(np.float64))
    assert abs(D.mad(g).item() - m) <= 1e-6 * max(m, 1e-30)


def test_median_f32():
    x = white_noise(100001, seed=3).astype(np.float32)
    g = torch.tensor(x, device="cuda")
    assert D.median(g).item() == O.median(x.astype(np.float64))
    assert abs(D.mad(g).item() - O.mad(x.astype(np.float64))) <= 1e-6 * O.mad(x.astype(np.float64))


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("ttype", [0, 1])
def test_univ_thresh_exact(filt, ttype):
    n = 1 << 14
    J = O.levels(n, 10 if ttype else 0, filt)
    x, _ = noisy_doppler(n, 4.0, seed=9)
    c = O.dwt(x, J, filt) if ttype == 0 else O.modwt_fwd(x, J, filt)
    g = _dev(c)
    for mn, mx in ((0, 0), (0, J - 1), (1, 3), (2, J - 1)):
        want = O.univ_thresh(c, n, J, ttype, mn, mx)
        got = D.univ_thresh(g, n, J, ttype, filt, mn, mx).item()
        assert got == want, (mn, mx)


@pytest.mark.parametrize("n", [4, 8, 1024, 1 << 17])
def test_interp_mse(n):
    rng = np.random.Generator(np.random.PCG64(n))
    xn, ye, yo = rng.standard_normal(n), rng.standard_normal(n // 2), rng.standard_normal(n // 2)
    want = O.interp_mse(xn, ye, yo)
    got = D.interp_mse(_dev(xn), _dev(ye), _dev(yo)).item()
    assert abs(got - want) <= 1e-13 * abs(want)
    again = D.interp_mse(_dev(xn), _dev(ye), _dev(yo)).item()
    assert again == got  # fixed-order reduction


SMOOTH_CASES = [
    # ttype, filt, threshtype, nlevels, hard, min, max, n, snr
    (0, 1, 0, 0, 1, 0, 6, 1024, 5.0),
    (0, 2, 1, 0, 1, 0, 7, 1024, 5.0),
    (0, 3, 1, 0, 0, 1, 6, 4096, 2.0),
    (0, 4, 1, 0, 1, 0, 7, 1 << 16, 5.0),
    (0, 1, 2, 0, 1, 0, 9, 1024, 5.0),
    (0, 2, 2, 0, 1, 0, 8, 4096, 5.0),
    (0, 4, 2, 0, 0, 2, 10, 1 << 14, 2.0),
    (1, 1, 0, 6, 1, 0, 5, 1024, 5.0),
    (1, 4, 1, 8, 1, 0, 7, 4096, 5.0),
    (1, 3, 1, 6, 0, 1, 5, 4096, 2.0),
    (1, 1, 2, 8, 1, 0, 7, 4096, 5.0),
    (1, 2, 2, 7, 0, 1, 6, 4096, 2.0),
    (1, 4, 2, 8, 1, 0, 7, 1 << 14, 5.0),
    (1, 3, 2, 8, 0, 0, 7, 1 << 14, 2.0),
]


@pytest.mark.parametrize("case", SMOOTH_CASES)
def test_smooth_parity(case):
    tt, f, th, nl, hard, mn, mx, n, snr = case
    x, _ = noisy_doppler(n, snr, seed=n + f + th)
    t0 = 0.3 if th == 0 else None
    want, t_want, rc = O.smooth(x, nl, tt, f, th, t0, hard, mn, mx, 0.01)
    assert rc == 0
    got, t_got = D.smooth(_dev(x), nl, tt, f, th, t0, hard, mn, mx, 0.01)
    assert same_t(t_got, t_want)
    assert O.relerr(got.cpu().numpy(), want) <= 1e-12


@pytest.mark.parametrize("tol", [0.0, 1e-6, 0.5])
@pytest.mark.parametrize("tt", [0, 1])
def test_cv_search_tolerances(tt, tol):
    """The device-side golden-section search against the oracle's host loop
    where the loop runs to its 51-iteration cap (tol 0), runs long (1e-6)
    and stops early (0.5); a second call repeats the threshold exactly."""
    n = 4096
    nl = 0 if tt == 0 else 8
    x, _ = noisy_doppler(n, 3.0, seed=23)
    want, t_want, rc = O.smooth(x, nl, tt, 4, 2, None, 0, 0, 7, tol)
    assert rc == 0
    got, t_got = D.smooth(_dev(x), nl, tt, 4, 2, None, 0, 0, 7, tol)
    assert same_t(t_got, t_want)
    assert O.relerr(got.cpu().numpy(), want) <= 1e-12
    _, t_again = D.smooth(_dev(x), nl, tt, 4, 2, None, 0, 0, 7, tol)
    assert t_again == t_got


# GPUSmooth configuration: MODWT J=10, Haar and LA8, levels 1..10, soft and
# hard, CV and universal, n = 2^18, SNR 5 and 2
@pytest.mark.parametrize("filt", [1, 4])
@pytest.mark.parametrize("threshtype", [1, 2])
@pytest.mark.parametrize("hard", [0, 1])
@pytest.mark.parametrize("snr", [5.0, 2.0])
def test_smooth_config_2p18(filt, threshtype, hard, snr):
    n = 1 << 18
    x, clean = noisy_doppler(n, snr, seed=17)
    want, t_want, rc = O.smooth(x, 10, 1, filt, threshtype, None, hard, 0, 9, 0.01)
    assert rc == 0
    got, t_got = D.smooth(_dev(x), 10, 1, filt, threshtype, None, hard, 0, 9, 0.01)
    assert same_t(t_got, t_want)
    g = got.cpu().numpy()
    assert O.relerr(g, want) <= 1e-12
    assert np.mean((g - clean) ** 2) < np.mean((x - clean) ** 2)


def test_smooth_f32_against_f64_oracle():
    n = 1 << 14
    x, _ = noisy_doppler(n, 5.0, seed=1)
    want, t_want, _ = O.smooth(x, 8, 1, 4, 1, None, 1, 0, 7, 0.01)
    got, t_got = D.smooth(_dev(x, torch.float32), 8, 1, 4, 1, None, 1, 0, 7, 0.01)
    assert abs(t_got - t_want) <= 1e-5 * t_want
    assert O.relerr(got.cpu().numpy(), want) <= 1e-5


def test_smooth_errors_leave_input():
    x = _dev(white_noise(256, seed=1))
    keep = x.clone()
    with pytest.raises(W.WaveHipError):
        D.smooth(x, 5, 1, 2, 2, None, 1, 2, 2, 0.01, inplace=True)  # CV with min == max
    with pytest.raises(W.WaveHipError):
        D.smooth(x, 5, 1, 2, 1, None, 1, 0, 5, 0.01, inplace=True)  # max >= nlevels
    assert torch.equal(x, keep)


# ------------------------------------------------ the .C entry points

_DP = C.POINTER(C.c_double)


def _ip(v):
    return C.byref(C.c_int(int(v)))


@pytest.mark.parametrize("filt", FILTERS)
def test_RgpuTransform_dwt_and_modwt(filt):
    L = W.FILTER_LENGTHS[filt]
    lib = W.lib()
    n = 2048
    x, _ = noisy_doppler(n, 5.0, seed=2)
    # DWT forward, in place on x as R's .C would see it
    J = O.levels(n, 0, filt)
    a = x.copy()
    lib.RgpuTransform(a.ctypes.data_as(_DP), None, _ip(n), _ip(1), _ip(J), _ip(0), _ip(filt), _ip(L))
    assert O.relerr(a, O.dwt(x, J, filt)) <= 1e-12
    lib.RgpuTransform(a.ctypes.data_as(_DP), None, _ip(n), _ip(0), _ip(J), _ip(0), _ip(filt), _ip(L))
    assert O.relerr(a, x) <= 1e-12
    # MODWT forward into xmod (R allocates 2 n log2(n); the tail stays zero)
    Jm = O.levels(n, 6, filt)
    xmod = np.zeros(2 * n * int(np.log2(n)))
    b = x.copy()
    lib.RgpuTransform(b.ctypes.data_as(_DP), xmod.ctypes.data_as(_DP), _ip(n), _ip(1), _ip(Jm),
                      _ip(1), _ip(filt), _ip(L))
    ref = O.modwt_fwd(x, Jm, filt)
    assert O.relerr(xmod[: 2 * n * Jm], ref) <= 1e-12
    assert not xmod[2 * n * Jm:].any()
    out = np.zeros(n)
    lib.RgpuTransform(out.ctypes.data_as(_DP), xmod.ctypes.data_as(_DP), _ip(n), _ip(0), _ip(Jm),
                      _ip(1), _ip(filt), _ip(L))
    assert O.relerr(out, x) <= 1e-12


@pytest.mark.parametrize("threshtype", [0, 1, 2])
def test_RgpuSmooth(threshtype):
    lib = W.lib()
    n, filt, J = 4096, 4, 8
    x, _ = noisy_doppler(n, 3.0, seed=8)
    y = x.copy()
    t = np.array([0.2])
    tol = np.array([0.01])
    lib.RgpuSmooth(y.ctypes.data_as(_DP), _ip(n), _ip(J), _ip(1), _ip(filt), _ip(8), _ip(threshtype),
                   t.ctypes.data_as(_DP), _ip(1), _ip(0), _ip(J - 1), tol.ctypes.data_as(_DP))
    want, t_want, _ = O.smooth(x, J, 1, filt, threshtype, 0.2 if threshtype == 0 else None, 1, 0,
                               J - 1, 0.01)
    assert O.relerr(y, want) <= 1e-12
    assert same_t(t[0], t_want)


def test_RgpuTransform_bad_args_untouched(capfd):
    lib = W.lib()
    x = np.arange(1000.0)
    keep = x.copy()
    lib.RgpuTransform(x.ctypes.data_as(_DP), None, _ip(1000), _ip(1), _ip(0), _ip(0), _ip(2), _ip(4))
    np.testing.assert_array_equal(x, keep)
    assert "Error in GPU transform" in capfd.readouterr().err
