"""The reference package's `.C` symbols (src/rfunctions.cuh:17-35; bound by
NAMESPACE:26-30) called through ctypes exactly as R's `.C` calls them:
pointers to host copies of every argument, outputs read back from those
copies.  Expected values are the golden vectors recorded through the
reference's own CPU glue (RcpuTransform / RcpuThreshold / RcpuSmooth,
src/rfunctions.cu:3-100, built from its sources by oracle/Makefile):
transforms and smoothed outputs within 1e-12 norm-wise (the GPU contracts
products into FMAs), thresholding of the same coefficients bit for bit,
chosen thresholds within 1e-12 relative.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import _lib as W
from wavecuda_amd import device as D
from wavecuda_amd.signals import batch_workload, noisy_doppler

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "reference_vectors.npz"
FLEN = {1: 2, 2: 4, 3: 6, 4: 8}
TOL = 1e-12


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ip(v):
    return C.byref(C.c_int(int(v)))


def transform(sym, x, xmod, n, sense, nlevels, ttype, f):
    """.C(sym, x=, xmod=, len=, sense=, nlevels=, ttype=, filter=, filterlen=)"""
    getattr(W.lib(), sym)(dp(x), dp(xmod) if xmod is not None else None, ip(n), ip(sense),
                          ip(nlevels), ip(ttype), ip(f), ip(FLEN[f]))


@pytest.mark.parametrize("sym", ["RcpuTransform", "RgpuTransform"])
def test_dwt_golden(gold, sym):
    keys = [k[:-2] for k in gold.files if k.startswith("dwt_") and k.endswith("_x")]
    for k in keys:
        f = int(k.split("_f")[1].split("_")[0])
        nl = int(k.split("_l")[1])
        x = gold[k + "_x"].copy()
        n = x.size
        transform(sym, x, None, n, 1, nl, 0, f)
        assert O.relerr(x, gold[k + "_fwd"]) <= TOL, k
        y = gold[k + "_fwd"].copy()
        transform(sym, y, None, n, 0, nl, 0, f)
        assert O.relerr(y, gold[k + "_inv"]) <= TOL, k
    x = gold["doppler1024_d4_x"].copy()
    transform(sym, x, None, 1024, 1, 0, 0, 2)
    assert O.relerr(x, gold["doppler1024_d4_fwd"]) <= TOL


@pytest.mark.parametrize("ttype,prefix", [(1, "modwt_"), (2, "modwtpo_")])
def test_modwt_golden(gold, ttype, prefix):
    keys = [k[:-2] for k in gold.files if k.startswith(prefix) and k.endswith("_x")]
    assert keys
    for k in keys:
        f = int(k.split("_f")[1].split("_")[0])
        x = gold[k + "_x"].copy()
        n = x.size
        want = gold[k + "_fwd"]
        J = want.size // (2 * n)
        # R allocates 2 n log2 n values (R/rfunctions.R:94); the tail stays 0
        xmod = np.zeros(2 * n * int(np.log2(n)))
        transform("RcpuTransform", x, xmod, n, 1, J, ttype, f)
        for b in range(2 * J):
            assert O.relerr(xmod[b * n:(b + 1) * n], want[b * n:(b + 1) * n]) <= TOL, (k, b)
        assert not np.any(xmod[2 * n * J:])
        xr = np.zeros(n)
        coef = want.copy()
        transform("RgpuTransform", xr, coef, n, 0, J, ttype, f)
        assert O.relerr(xr, gold[k + "_inv"]) <= TOL, k
        np.testing.assert_array_equal(coef, want)  # the inverse leaves xmod as it was


def test_threshold_golden(gold):
    keys = [k for k in gold.files if k.startswith("thr_") and k.endswith("_out")]
    assert len(keys) >= 30
    for k in keys:
        tt = int(k.split("_t")[1].split("_")[0])
        f = int(k.split("_f")[1].split("_")[0])
        hard = int(k.split("_h")[1].split("_")[0])
        mn, mx = (int(v) for v in k.split("_m")[1].split("_")[:2])
        c = gold[f"thr_t{tt}_f{f}_in"].copy()
        t = np.array([gold[f"thr_t{tt}_f{f}_t"][0]])
        n = 128
        J = O.levels(n, 0, f)
        if tt == 0:
            W.lib().RcpuThreshold(dp(c), None, ip(n), ip(J), ip(0), ip(f), ip(FLEN[f]), dp(t),
                                  ip(hard), ip(mn), ip(mx))
        else:
            x = np.zeros(n)
            W.lib().RcpuThreshold(dp(x), dp(c), ip(n), ip(J), ip(1), ip(f), ip(FLEN[f]), dp(t),
                                  ip(hard), ip(mn), ip(mx))
        np.testing.assert_array_equal(c, gold[k], err_msg=k)


def test_threshold_packet_order():
    """Packet-ordered MODWT details sit at odd positions of each level
    (src/thresh.cpp:56-63): thresholding them equals converting to time
    order, thresholding the D blocks and converting back."""
    n, J, f = 256, 4, 2
    x, _ = noisy_doppler(n, 3.0, seed=21)
    po = O.modwt_po_fwd(x, J, f)
    t = np.array([0.3])
    got = po.copy()
    W.lib().RcpuThreshold(dp(np.zeros(n)), dp(got), ip(n), ip(J), ip(2), ip(f), ip(FLEN[f]),
                          dp(t), ip(1), ip(1), ip(3))
    want = po.copy()
    for l in range(1, 4):
        blk = want[2 * n * l: 2 * n * (l + 1)]
        blk[1::2] = [O.oracle().wvo_thresh_coef(v, 0.3, 1) for v in blk[1::2]]
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("sym", ["RcpuSmooth", "RgpuSmooth"])
def test_smooth_golden(gold, sym):
    idx = sorted({int(k.split("_")[1]) for k in gold.files if k.startswith("smooth_")})
    assert len(idx) >= 20
    for i in idx:
        tt, f, th, nl, hard, mn, mx, n = (int(v) for v in gold[f"smooth_{i}_args"])
        snr = float(gold[f"smooth_{i}_snr"][0])
        x, _ = noisy_doppler(n, snr, seed=100 + i)
        t = np.array([0.25 if th == 0 else 0.0])
        tol = np.array([0.01])
        getattr(W.lib(), sym)(dp(x), ip(n), ip(nl), ip(tt), ip(f), ip(FLEN[f]), ip(th), dp(t),
                              ip(hard), ip(mn), ip(mx), dp(tol))
        assert O.relerr(x, gold[f"smooth_{i}_y"]) <= TOL, i
        want_t = float(gold[f"smooth_{i}_t"][0])
        assert abs(t[0] - want_t) <= TOL * abs(want_t), (i, t[0], want_t)


def test_errors_leave_outputs_untouched(capfd):
    x = np.arange(1000.0)
    keep = x.copy()
    transform("RcpuTransform", x, None, 1000, 1, 0, 0, 2)  # not a power of two
    np.testing.assert_array_equal(x, keep)
    t = np.array([0.5])
    W.lib().RcpuThreshold(dp(x[:64].copy()), None, ip(64), ip(3), ip(0), ip(2), ip(4), dp(t),
                          ip(0), ip(2), ip(5))  # maxlevel beyond the levels
    W.lib().RcpuSmooth(dp(x), ip(1000), ip(0), ip(0), ip(2), ip(4), ip(1), dp(t), ip(0), ip(0),
                       ip(3), dp(np.array([0.01])))
    np.testing.assert_array_equal(x, keep)
    err = capfd.readouterr().err
    assert err.count("Error in GPU") == 3


@pytest.mark.parametrize("n,J", [(1 << 22, 10), (1 << 20, 5)])
def test_large_transfer_matches_device_path(n, J):
    """LA8 MODWT through the staging ring of 32 MiB slots, identical to the
    device-pointer path bit for bit: J = 10 at n = 2^22 (the 640 MiB
    coefficient array crosses in 20 full chunks each way, x in one), J = 5
    at n = 2^20 (80 MiB: two full chunks and a half one; x in a partial
    first chunk)."""
    f = 4
    x = np.random.Generator(np.random.PCG64(5)).standard_normal(n)
    xmod = np.zeros(2 * n * J)
    transform("RcpuTransform", x.copy(), xmod, n, 1, J, 1, f)
    xd = torch.from_numpy(x).cuda()
    want = D.modwt_fwd(xd, J, f).cpu().numpy()
    np.testing.assert_array_equal(xmod, want)
    xr = np.zeros(n)
    transform("RgpuTransform", xr, xmod, n, 0, J, 1, f)
    np.testing.assert_array_equal(xr, D.modwt_inv(torch.from_numpy(want).cuda(), n, J, f)
                                  .cpu().numpy())


def test_smooth_list_equals_single_signal_calls():
    """RgpuSmoothList over the mixed batch workload (a leading sample of it)
    gives, signal by signal, exactly what RgpuSmooth with the universal
    threshold over every level gives; a second call with the same table
    reuses the plan."""
    x, lens, filters, ttypes, levels = batch_workload()
    m = 256
    lens, filters, ttypes, levels = (np.ascontiguousarray(a[:m], dtype=np.int32)
                                     for a in (lens, filters, ttypes, levels))
    off = np.concatenate([[0], np.cumsum(lens.astype(np.int64))])
    xs = np.ascontiguousarray(x[: off[-1]], dtype=np.float64)
    ipa = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    for rep in range(2):
        y = xs.copy()
        th = np.zeros(m)
        W.lib().RgpuSmoothList(dp(y), ip(m), ipa(lens), ipa(filters), ipa(ttypes), ipa(levels),
                               ip(1), dp(th))
        for i in (0, 1, 17, 100, 255):
            n = int(lens[i])
            J = O.levels(n, int(levels[i]), int(filters[i]))
            xi = xs[off[i]:off[i + 1]].copy()
            t = np.array([0.0])
            W.lib().RgpuSmooth(dp(xi), ip(n), ip(J), ip(ttypes[i]), ip(filters[i]),
                               ip(FLEN[int(filters[i])]), ip(1), dp(t), ip(1), ip(0),
                               ip(J - 1), dp(np.array([0.01])))
            np.testing.assert_array_equal(y[off[i]:off[i + 1]], xi, err_msg=str((rep, i)))
            assert th[i] == t[0]
