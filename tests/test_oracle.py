"""Pins of the CPU oracle (oracle/wvoracle.c), CPU only.

1. Golden vectors recorded from a build of the reference's own CPU sources
   (tests/golden/reference_vectors.npz): the oracle must reproduce them
   bit for bit (same term order, no FMA contraction in either build).
2. The live reference build (oracle/_ref), when present, on fresh inputs.
3. What the mathematics fixes, for the parts the reference lacks (C6 / LA8
   MODWT) or that a mistake could slip past: filter moments and
   orthonormality, impulse known answers of the convolution alignment,
   orthogonality of the dense DWT matrix, the MODWT decimation identity
   against the DWT, the per-level MODWT energy law, the MODWT inverse as
   half the adjoint of the forward level, brute-force medians and the
   threshold rule.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_bind as O
from wavecuda_amd.signals import noisy_doppler

GOLD = Path(__file__).resolve().parent / "golden" / "reference_vectors.npz"
FLEN = {1: 2, 2: 4, 3: 6, 4: 8}


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def coeffs(f):
    c = np.zeros(8)
    O.oracle().wvo_filter_coeffs(f, O._p(c))
    return c[: FLEN[f]]


def gfilt(c):
    L = len(c)
    return np.array([(-1) ** k * c[L - 1 - k] for k in range(L)])


# ------------------------------------------------------------ golden vectors

def test_golden_dwt(gold):
    keys = [k[:-2] for k in gold.files if k.startswith("dwt_") and k.endswith("_x")]
    assert len(keys) > 40
    for k in keys:
        f = int(k.split("_f")[1].split("_")[0])
        nl = int(k.split("_l")[1])
        x = gold[k + "_x"]
        y = O.dwt(x, nl, f, 1)
        np.testing.assert_array_equal(y, gold[k + "_fwd"], err_msg=k)
        z = O.dwt(y, nl, f, 0)
        np.testing.assert_array_equal(z, gold[k + "_inv"], err_msg=k)


def test_golden_doppler_d4_1024(gold):
    x = gold["doppler1024_d4_x"]
    np.testing.assert_array_equal(x, noisy_doppler(1024, 5.0, seed=10)[0])
    assert O.levels(1024, 0, 2) == 9  # the reference gives 9 levels for D4 at n=1024
    y = O.dwt(x, 0, 2, 1)
    np.testing.assert_array_equal(y, gold["doppler1024_d4_fwd"])
    np.testing.assert_array_equal(O.dwt(y, 0, 2, 0), gold["doppler1024_d4_inv"])


def test_golden_modwt(gold):
    keys = [k[:-2] for k in gold.files if k.startswith("modwt_") and k.endswith("_x")]
    assert len(keys) > 15
    for k in keys:
        f = int(k.split("_f")[1].split("_")[0])
        n = int(k.split("_n")[1].split("_")[0])
        nl = int(k.split("_l")[1])
        J = O.levels(n, nl, f)
        x = gold[k + "_x"]
        xm = O.modwt_fwd(x, J, f)
        np.testing.assert_array_equal(xm, gold[k + "_fwd"], err_msg=k)
        # the inverse also overwrites the S_{j-1} slots, as the reference does
        xm2 = xm.copy()
        xr = np.zeros(n)
        O.oracle().wvo_modwt(O._p(xr), O._p(xm2), n, 0, J, f)
        np.testing.assert_array_equal(xr, gold[k + "_inv"], err_msg=k)
        np.testing.assert_array_equal(xm2, gold[k + "_invxmod"], err_msg=k)


def test_golden_threshold(gold):
    outs = [k for k in gold.files if k.startswith("thr_") and k.endswith("_out")]
    assert len(outs) >= 30
    for k in outs:
        parts = k.split("_")
        ttype, f = int(parts[1][1:]), int(parts[2][1:])
        hard = int(parts[3][1:])
        mn, mx = int(parts[4][1:]), int(parts[5])
        c = gold[f"thr_t{ttype}_f{f}_in"]
        t = float(gold[f"thr_t{ttype}_f{f}_t"][0])
        n = 128
        J = O.levels(n, 0, f)
        got = O.threshold(c, n, J, ttype, f, t, hard, mn, mx)
        np.testing.assert_array_equal(got, gold[k], err_msg=k)


def test_golden_median_mad(gold):
    for i in range(7):
        x = gold[f"median_{i}_x"]
        assert O.median(x) == gold[f"median_{i}_med"][0]
        assert O.mad(x) == gold[f"median_{i}_mad"][0]


def test_golden_univ(gold):
    us = [k for k in gold.files if k.startswith("univ_") and k.endswith("_u")]
    assert len(us) == 16
    for k in us:
        parts = k.split("_")
        ttype, f = int(parts[1][1:]), int(parts[2][1:])
        mn, mx = int(parts[3][1:]), int(parts[4])
        c = gold[f"univ_t{ttype}_f{f}_coef"]
        n = 256
        J = O.levels(n, 0, f)
        assert O.univ_thresh(c, n, J, ttype, mn, mx) == gold[k][0], k


def test_golden_interp_mse(gold):
    for n in (4, 8, 1024):
        m = O.interp_mse(gold[f"imse_n{n}_xn"], gold[f"imse_n{n}_ye"], gold[f"imse_n{n}_yo"])
        assert m == gold[f"imse_n{n}_m"][0]


def test_golden_smooth(gold):
    idx = sorted({int(k.split("_")[1]) for k in gold.files if k.startswith("smooth_")})
    assert len(idx) >= 20
    for i in idx:
        tt, f, th, nl, hard, mn, mx, n = (int(v) for v in gold[f"smooth_{i}_args"])
        snr = float(gold[f"smooth_{i}_snr"][0])
        x, _ = noisy_doppler(n, snr, seed=100 + i)
        t0 = 0.25 if th == 0 else None
        y, t, rc = O.smooth(x, nl, tt, f, th, t0, hard, mn, mx, 0.01)
        assert rc == 0
        np.testing.assert_array_equal(y, gold[f"smooth_{i}_y"], err_msg=f"smooth_{i}")
        assert t == gold[f"smooth_{i}_t"][0]


# ------------------------------------------------- live reference build

@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build not present")
def test_live_reference_random():
    import ctypes as C
    R = O.ref()
    rng = np.random.Generator(np.random.PCG64(77))
    for f in (1, 2, 3, 4):
        for n in (FLEN[f], 32, 512, 4096):
            x = rng.standard_normal(n)
            a = x.copy()
            b = x.copy()
            O.oracle().wvo_dwt(O._p(a), n, 1, 0, f)
            R.ref_transform(O._p(b), None, n, 1, 0, 0, f, FLEN[f])
            np.testing.assert_array_equal(a, b)
    for f in (1, 2):
        n, J = 2048, 6
        x = rng.standard_normal(n)
        xm = O.modwt_fwd(x, J, f)
        ym = np.zeros(2 * n * J)
        R.ref_transform(O._p(x.copy()), O._p(ym), n, 1, J, 1, f, FLEN[f])
        np.testing.assert_array_equal(xm, ym)
    x = rng.standard_normal(999)
    assert O.median(x) == R.ref_median(O._p(x.copy()), 999)
    del C


# ------------------------------------------------------------- mathematics

@pytest.mark.parametrize("f", [1, 2, 3, 4])
def test_filter_orthonormal_and_moments(f):
    c = coeffs(f)
    L = len(c)
    assert abs(np.sum(c) - np.sqrt(2)) < 1e-14
    assert abs(np.sum(c * c) - 1) < 1e-14
    for m in range(1, L // 2):
        assert abs(np.sum(c[: L - 2 * m] * c[2 * m:])) < 1e-14
    # vanishing moments of the wavelet filter: Haar 1, D4 2, C6 2, LA8 4
    g = gfilt(c)
    k = np.arange(L)
    nmom = {1: 1, 2: 2, 3: 2, 4: 4}[f]
    for pw in range(nmom):
        assert abs(np.sum(g * k ** pw)) < 1e-12 * max(1, L ** pw)


def test_filter_closed_forms():
    s3, s2, s7 = np.sqrt(3), np.sqrt(2), np.sqrt(7)
    d4 = np.array([1 + s3, 3 + s3, 3 - s3, 1 - s3]) / (4 * s2)
    np.testing.assert_allclose(coeffs(2), d4, rtol=0, atol=1e-15)
    c6 = s2 * np.array([(1 - s7) / 32, (5 + s7) / 32, (7 + s7) / 16, (7 - s7) / 16,
                        (1 - s7) / 32, (-3 + s7) / 32])
    np.testing.assert_allclose(coeffs(3), c6, rtol=0, atol=1e-15)


def dense_dwt(n, J, f):
    W = np.zeros((n, n))
    for p in range(n):
        e = np.zeros(n)
        e[p] = 1
        W[:, p] = O.dwt(e, J, f, 1)
    return W


@pytest.mark.parametrize("f", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [8, 16, 64])
def test_dwt_dense_orthogonal(f, n):
    J = O.levels(n, 0, f)
    W = dense_dwt(n, J, f)
    np.testing.assert_allclose(W @ W.T, np.eye(n), atol=1e-14)
    # the inverse is the transpose
    for p in range(0, n, 3):
        e = np.zeros(n)
        e[p] = 1
        np.testing.assert_allclose(O.dwt(e, J, f, 0), W.T[:, p], atol=1e-15)


@pytest.mark.parametrize("f", [1, 2, 3, 4])
def test_dwt_impulse_alignment(f):
    """Level 1 of x = e_p: s[k] = C_{(p-2k) mod n}, d[k] = g_{(p-2k) mod n}
    (zero outside 0..L-1) -- the convolution alignment."""
    n = 32
    c = coeffs(f)
    g = gfilt(c)
    L = len(c)
    for p in range(n):
        e = np.zeros(n)
        e[p] = 1.0
        y = O.dwt(e, 1, f, 1)
        for k in range(n // 2):
            r = (p - 2 * k) % n
            es = c[r] if r < L else 0.0
            ed = g[r] if r < L else 0.0
            assert abs(y[2 * k] - es) < 1e-15 and abs(y[2 * k + 1] - ed) < 1e-15


@pytest.mark.parametrize("f", [1, 2, 3, 4])
def test_dwt_parseval_constant(f):
    n = 4096
    x, _ = noisy_doppler(n, 5.0, seed=1)
    J = O.levels(n, 0, f)
    y = O.dwt(x, J, f)
    assert abs(np.sum(y * y) - np.sum(x * x)) < 1e-10 * np.sum(x * x)
    c = np.full(n, 1.7)
    yc = O.dwt(c, J, f)
    det = np.ones(n, bool)
    det[:: 1 << J] = False
    assert np.max(np.abs(yc[det])) < 1e-12
    np.testing.assert_allclose(yc[:: 1 << J], 1.7 * 2 ** (J / 2), rtol=1e-13)


@pytest.mark.parametrize("f", [1, 2, 3, 4])
def test_modwt_decimation_identity(f):
    """S_j[m 2^j] = s_j[m] and D_j[m 2^j] = d_j[m]: the MODWT contains the DWT."""
    n = 1024
    J = O.levels(n, 0, f)
    x, _ = noisy_doppler(n, 4.0, seed=2)
    xm = O.modwt_fwd(x, J, f)
    y = O.dwt(x, J, f)
    for j in range(1, J + 1):
        S = xm[2 * n * (j - 1): 2 * n * (j - 1) + n]
        D = xm[2 * n * (j - 1) + n: 2 * n * j]
        m = np.arange(n >> j)
        np.testing.assert_allclose(D[m << j], y[(1 << (j - 1)) + (m << j)], atol=1e-13)
        if j == J:
            np.testing.assert_allclose(S[m << j], y[m << J], atol=1e-13)


@pytest.mark.parametrize("f", [1, 2, 3, 4])
def test_modwt_energy_and_pr(f):
    n = 2048
    J = O.levels(n, 8, f)
    x, _ = noisy_doppler(n, 3.0, seed=4)
    xm = O.modwt_fwd(x, J, f)
    prev = np.sum(x * x)
    for j in range(1, J + 1):
        S = xm[2 * n * (j - 1): 2 * n * (j - 1) + n]
        D = xm[2 * n * (j - 1) + n: 2 * n * j]
        e = np.sum(S * S) + np.sum(D * D)
        assert abs(e - 2 * prev) < 1e-11 * e
        prev = np.sum(S * S)
    np.testing.assert_allclose(O.modwt_inv(xm, n, J, f), x, atol=1e-13)


@pytest.mark.parametrize("f", [1, 2, 3, 4])
def test_modwt_inverse_is_half_adjoint(f):
    """One level: inverse = 1/2 (forward)^T, checked on dense matrices."""
    n = 16
    A = np.zeros((2 * n, n))
    for p in range(n):
        e = np.zeros(n)
        e[p] = 1
        A[:, p] = O.modwt_fwd(e, 1, f)
    B = np.zeros((n, 2 * n))
    for q in range(2 * n):
        e = np.zeros(2 * n)
        e[q] = 1
        B[:, q] = O.modwt_inv(e, n, 1, f)
    np.testing.assert_allclose(B, 0.5 * A.T, atol=1e-15)


def test_median_mad_bruteforce():
    rng = np.random.Generator(np.random.PCG64(3))
    for n in (1, 2, 5, 64, 1001, 4096):
        x = rng.standard_normal(n)
        assert O.median(x) == np.median(x)
        assert O.mad(x) == 1.4826 * np.median(np.abs(x - np.median(x)))


def test_threshold_rule():
    for c, t in ((0.5, 0.5), (-0.5, 0.5), (0.49, 0.5), (2.0, 0.5), (-2.0, 0.5)):
        h = O.oracle().wvo_thresh_coef(c, t, 0)
        s = O.oracle().wvo_thresh_coef(c, t, 1)
        assert h == (0.0 if abs(c) < t else c)
        assert s == (0.0 if abs(c) < t else np.sign(c) * (abs(c) - t))


def test_level_rules():
    assert O.levels(1024, 0, 1) == 10
    assert O.levels(1024, 0, 2) == 9
    assert O.levels(1024, 0, 3) == 8
    assert O.levels(1024, 0, 4) == 8
    assert O.levels(256, 0, 4) == 6
    assert O.levels(256, 7, 4) == 0
    assert O.levels(1000, 0, 2) == 0
    assert O.oracle().wvo_check_len_levels_mm(1024, 5, 0, 4, 4) == 5
    assert O.oracle().wvo_check_len_levels_mm(1024, 5, 0, 5, 4) == 0


def test_cv_rejects_equal_levels():
    x, _ = noisy_doppler(256, 5.0, seed=1)
    t, y, tr, st = O.cvt(x, 5, 1, 2, 1, 0.01, 2, 2)
    assert st != 0
    np.testing.assert_array_equal(y, x)


@pytest.mark.parametrize("f", [1, 2, 3, 4])
def test_modwt_shift_equivariance(f):
    """Every MODWT position is a DWT coefficient of a shifted signal: with
    y = x shifted left by r, S_j^x[m 2^j + r] = s_j^y[m] and
    D_j^x[m 2^j + r] = d_j^y[m].  Together with the golden-pinned DWT this
    pins the C6 / LA8 MODWT (absent from the reference) at every position,
    not only at the decimated ones; the term order of the two recursions is
    the same, so the values agree bit for bit."""
    n = 256
    J = O.levels(n, 6, f)
    x, _ = noisy_doppler(n, 4.0, seed=7)
    xm = O.modwt_fwd(x, J, f)
    for r in range(1 << J):
        y = O.dwt(np.roll(x, -r), J, f)
        for j in range(1, J + 1):
            S = xm[2 * n * (j - 1): 2 * n * (j - 1) + n]
            D = xm[2 * n * (j - 1) + n: 2 * n * j]
            m = np.arange(n >> j)
            if (r >> j) == 0:  # r < 2^j: position m 2^j + r of level j
                np.testing.assert_array_equal(D[(m << j) + r], y[(1 << (j - 1)) + (m << j)])
                if j == J:
                    np.testing.assert_array_equal(S[(m << j) + r], y[m << J])
    # and one shift of the input shifts every block by one
    xs = O.modwt_fwd(np.roll(x, 1), J, f)
    for b in range(2 * J):
        np.testing.assert_array_equal(xs[b * n:(b + 1) * n], np.roll(xm[b * n:(b + 1) * n], 1))


@pytest.mark.parametrize("f", [1, 2, 3, 4])
@pytest.mark.parametrize("n,nl", [(64, 0), (4096, 0), (1 << 16, 10), (1 << 16, 0)])
def test_openmp_variants_equal_serial(f, n, nl):
    """The OpenMP restatements (the all-cores CPU baseline) compute every
    value with the same operations as the serial oracle: bit-identical, DWT
    and MODWT, forward and inverse."""
    x, _ = noisy_doppler(n, 5.0, seed=n + f)
    J = O.levels(n, nl, f)
    if J == 0:
        return
    yd = O.dwt(x, J, f, 1)
    np.testing.assert_array_equal(O.dwt(x, J, f, 1, omp=True), yd)
    np.testing.assert_array_equal(O.dwt(yd, J, f, 0, omp=True), O.dwt(yd, J, f, 0))
    Jm = min(J, 10)
    xm = O.modwt_fwd(x, Jm, f)
    np.testing.assert_array_equal(O.modwt_fwd(x, Jm, f, omp=True), xm)
    rng = np.random.Generator(np.random.PCG64(n))
    w = rng.standard_normal(xm.size)  # inconsistent coefficients: a left inverse is not pinned by PR
    np.testing.assert_array_equal(O.modwt_inv(w, n, Jm, f, omp=True), O.modwt_inv(w, n, Jm, f))
