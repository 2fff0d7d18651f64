"""GPU parity of the transforms against the CPU oracle (through the C ABI).

Bar (north star): FP64 results within 1e-12 norm-wise relative
(max|a-b| / max|b|, per level block) of the oracle; FP32 within 1e-5 of
the FP64 oracle; perfect reconstruction to the same bounds.
"""
import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import device as D
from wavecuda_amd.signals import noisy_doppler, white_noise

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5
FILTERS = [1, 2, 3, 4]


def _dev(a, dtype=torch.float64):
    return torch.tensor(a, dtype=dtype, device="cuda")


def _blocks_relerr(got, ref, n):
    """Max over the n-long blocks of the norm-wise relative error."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    return max(O.relerr(got[i:i + n], ref[i:i + n]) for i in range(0, len(ref), n))


def _dwt_level_relerr(got, ref, J):
    """Max over DWT levels (and the coarse scaling block) of the error."""
    n = len(ref)
    errs = []
    for j in range(1, J + 1):
        idx = np.arange(1 << (j - 1), n, 1 << j)
        errs.append(O.relerr(got[idx], ref[idx]))
    idx = np.arange(0, n, 1 << J)
    errs.append(O.relerr(got[idx], ref[idx]))
    return max(errs)


MODWT_CASES = [(8, 0), (16, 0), (64, 2), (256, 0), (1024, 5), (4096, 0), (1 << 14, 10),
               (1 << 16, 0), (1 << 18, 10)]


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("n,nlev", MODWT_CASES)
def test_modwt_parity_f64(filt, n, nlev):
    x = white_noise(n, seed=n + filt)
    if O.levels(n, nlev, filt) == 0:
        # n shorter than the filter allows (reference src/wavutils.cpp:48-65):
        # the library refuses, as the reference's level check does
        with pytest.raises(ValueError):
            D.modwt_fwd(_dev(x), nlev, filt)
        return
    J = O.levels(n, nlev, filt)
    ref = O.modwt_fwd(x, J, filt)
    got = D.modwt_fwd(_dev(x), J, filt)
    torch.cuda.synchronize()
    assert _blocks_relerr(got.cpu().numpy(), ref, n) <= TOL64
    # inverse of the oracle's coefficients
    xr_ref = O.modwt_inv(ref, n, J, filt)
    xr = D.modwt_inv(_dev(ref), n, J, filt).cpu().numpy()
    assert O.relerr(xr, xr_ref) <= TOL64
    assert O.relerr(xr, x) <= TOL64


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("n,nlev", MODWT_CASES[2:])
def test_modwt_parity_f32(filt, n, nlev):
    J = O.levels(n, nlev, filt)
    x, _ = noisy_doppler(n, 5.0, seed=3)
    ref = O.modwt_fwd(x, J, filt)
    got = D.modwt_fwd(_dev(x, torch.float32), J, filt).cpu().numpy()
    assert _blocks_relerr(got, ref, n) <= TOL32
    xr = D.modwt_inv(_dev(ref, torch.float32), n, J, filt).cpu().numpy()
    assert O.relerr(xr, x) <= TOL32


# (4096, 1..3): one tile pass of 1..3 levels over the signal itself;
# (2^17, 6..8): a five-level tile pass, then a last tile pass of 1..3
# levels (the inverse's in-place top s for one and two levels)
DWT_CASES = [(2, 0), (4, 0), (8, 0), (16, 0), (64, 0), (1024, 0), (1024, 3), (4096, 1),
             (4096, 2), (4096, 3), (8192, 0), (16384, 0), (16384, 7), (1 << 16, 0),
             (1 << 16, 4), (1 << 17, 6), (1 << 17, 7), (1 << 17, 8), (1 << 18, 0),
             (1 << 20, 12)]


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("n,nlev", DWT_CASES)
def test_dwt_parity_f64(filt, n, nlev):
    x, _ = noisy_doppler(n, 5.0, seed=n)
    if O.levels(n, nlev, filt) == 0:
        with pytest.raises(ValueError):
            D.dwt_fwd(_dev(x), nlev, filt)
        return
    J = O.levels(n, nlev, filt)
    ref = O.dwt(x, J, filt, sense=1)
    got = D.dwt_fwd(_dev(x), J, filt).cpu().numpy()
    assert _dwt_level_relerr(got, ref, J) <= TOL64
    xr_ref = O.dwt(ref, J, filt, sense=0)
    xr = D.dwt_inv(_dev(ref), J, filt).cpu().numpy()
    assert O.relerr(xr, xr_ref) <= TOL64
    assert O.relerr(xr, x) <= TOL64


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("n,nlev", [(4096, 1), (4096, 3), (1 << 16, 4), (1 << 17, 6),
                                    (1 << 17, 7), (1 << 18, 0)])
@pytest.mark.parametrize("mode", ["soft_all", "hard_range", "levels"])
def test_dwt_inverse_thresholded(filt, n, nlev, mode):
    """The threshold fused into every inverse pass (each pass reads its
    levels' details out of its staged span, or in place for levels 1 and
    2) against the oracle's threshold followed by its inverse."""
    J = O.levels(n, nlev, filt)
    x, _ = noisy_doppler(n, 3.0, seed=n + filt)
    c = O.dwt(x, J, filt, sense=1)
    if mode == "levels":
        lams = np.linspace(0.05, 0.6, J)
        ref = O.threshold_levels_(c.copy(), n, J, 0, filt, lams, 1)
        thr = {"hardness": 1, "minlevel": 0, "maxlevel": J - 1,
               "levels": torch.tensor(lams, device="cuda")}
    else:
        hard = 0 if mode == "hard_range" else 1
        lo, hi = (1, J - 1) if mode == "hard_range" and J > 1 else (0, J - 1)
        ref = O.threshold(c, n, J, 0, filt, 0.3, hard, lo, hi)
        thr = {"value": 0.3, "hardness": hard, "minlevel": lo, "maxlevel": hi}
    xr_ref = O.dwt(ref, J, filt, sense=0)
    xr = D.dwt_inv(_dev(c), J, filt, thresh=thr).cpu().numpy()
    assert O.relerr(xr, xr_ref) <= TOL64


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("n,nlev", [(1024, 0), (4096, 2), (16384, 0), (1 << 17, 6),
                                    (1 << 17, 7), (1 << 18, 0)])
def test_dwt_parity_f32(filt, n, nlev):
    J = O.levels(n, nlev, filt)
    x = white_noise(n, seed=7)
    ref = O.dwt(x, J, filt, sense=1)
    got = D.dwt_fwd(_dev(x, torch.float32), J, filt).cpu().numpy()
    assert _dwt_level_relerr(got, ref, J) <= TOL32
    xr = D.dwt_inv(_dev(ref, torch.float32), J, filt).cpu().numpy()
    assert O.relerr(xr, x) <= TOL32


@pytest.mark.parametrize("filt", FILTERS)
def test_dwt_in_place(filt):
    n = 1 << 16
    x = white_noise(n, seed=11)
    J = O.levels(n, 0, filt)
    t = _dev(x)
    D.dwt_fwd(t, J, filt, out=t)
    ref = O.dwt(x, J, filt)
    assert O.relerr(t.cpu().numpy(), ref) <= TOL64
    D.dwt_inv(t, J, filt, out=t)
    assert O.relerr(t.cpu().numpy(), x) <= TOL64


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("kind", ["dwt", "modwt"])
def test_large_perfect_reconstruction(filt, kind):
    """n = 2^24 (the benchmark size): size-independent PR property."""
    n = 1 << 24
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(5))
    if kind == "dwt":
        y = D.dwt_fwd(x, 0, filt)
        xr = D.dwt_inv(y, 0, filt)
    else:
        y = D.modwt_fwd(x, 10, filt)
        xr = D.modwt_inv(y, n, 10, filt)
        del y
    err = (xr - x).abs().max().item() / x.abs().max().item()
    assert err <= TOL64
