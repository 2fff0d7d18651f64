"""GPU checks of the many-signal MODWT, the adjoint kernels and the
differentiable pipeline.

* Batched transforms equal the single-signal ones bit for bit.
* Dense matrices: the forward / inverse transform matrices are built by
  the CPU oracle from basis vectors; each adjoint kernel applied to basis
  vectors must give the oracle matrix's transpose (<= 1e-12 of its largest
  entry) -- the adjoints are checked, not assumed equal to inverses.
* <F x, y> = <x, F^T y> at n = 2^14 (relative 1e-12, FP64).
* Central finite differences (h = 1e-6, FP64) of a scalar loss through
  forward -> per-level soft threshold -> inverse agree with the autograd
  gradient in x and in every level's threshold to 1e-6.
* Fitting per-level thresholds by Adam lowers the error below the
  universal threshold's.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_bind as O
from wavecuda_amd import _lib as W
from wavecuda_amd import autograd as A
from wavecuda_amd import device as D
from wavecuda_amd.signals import noisy_doppler, white_noise

pytestmark = pytest.mark.gpu
FILTERS = [1, 2, 3, 4]


def _cuda(a, dtype=torch.float64):
    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda")


def _close(a, b, tol):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b))) <= tol * max(1.0, float(np.max(np.abs(b))))


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_many_equals_single(filt, dtype):
    n, B = 1 << 12, 5
    J = O.levels(n, 8, filt)
    x = torch.stack([_cuda(white_noise(n, seed=s), dtype) for s in range(B)])
    w = A.ModwtFn.apply(x, J, filt)
    xr = A.ImodwtFn.apply(w, n, J, filt)
    for b in range(B):
        wb = D.modwt_fwd(x[b].contiguous(), J, filt)
        assert torch.equal(w[b], wb), f"signal {b}"
        assert torch.equal(xr[b], D.modwt_inv(wb, n, J, filt)), f"signal {b}"


@pytest.mark.parametrize("filt", FILTERS)
def test_modwt_adjoints_dense(filt):
    n = 256 if filt != 1 else 64
    J = min(O.levels(n, 0, filt), 6)
    m = 2 * n * J
    eye_n = np.eye(n)
    F = np.stack([O.modwt_fwd(eye_n[i], J, filt) for i in range(n)], axis=1)  # (m, n)
    eye_m = np.eye(m)
    R = np.stack([O.modwt_inv(eye_m[k], n, J, filt) for k in range(m)], axis=1)  # (n, m)
    # adjoint of the forward on every basis vector of the coefficient space
    gx = torch.empty(m, n, dtype=torch.float64, device="cuda")
    W.check(W.lib().wh_modwt_fwd_adj(_cuda(eye_m).data_ptr(), gx.data_ptr(), n, m, J, filt,
                                     W.WH_F64, None), "fwd_adj")
    assert _close(gx.cpu().numpy(), F, 1e-12)  # row k = F^T e_k = F[k, :]
    gw = torch.empty(n, m, dtype=torch.float64, device="cuda")
    W.check(W.lib().wh_modwt_inv_adj(_cuda(eye_n).data_ptr(), gw.data_ptr(), n, n, J, filt,
                                     W.WH_F64, None), "inv_adj")
    assert _close(gw.cpu().numpy(), R, 1e-12)  # row i = R^T e_i = R[i, :]
    # the inverse is not half the adjoint's transpose by assumption: R F = I
    assert _close(R @ F, eye_n, 1e-12)


@pytest.mark.parametrize("filt", FILTERS)
def test_dwt_adjoints_dense(filt):
    n = 64
    J = O.levels(n, 0, filt)
    eye = np.eye(n)
    Wf = np.stack([O.dwt(eye[i], J, filt) for i in range(n)], axis=1)
    Wi = np.stack([O.dwt(eye[i], J, filt, sense=0) for i in range(n)], axis=1)

    def rows(fn):
        out = []
        for k in range(n):
            src = _cuda(eye[k])
            dst = torch.empty_like(src)
            W.check(fn(src.data_ptr(), dst.data_ptr(), n, J, filt, W.WH_F64, None))
            out.append(dst.cpu().numpy())
        return np.stack(out)

    assert _close(rows(W.lib().wh_dwt_fwd_adj), Wf, 1e-12)  # row k = W^T e_k = W[k, :]
    assert _close(rows(W.lib().wh_dwt_inv_adj), Wi, 1e-12)


@pytest.mark.parametrize("filt", FILTERS)
def test_dot_product_identities(filt):
    n, B = 1 << 14, 3
    J = O.levels(n, 10, filt)
    rng = np.random.Generator(np.random.PCG64(filt))
    x = _cuda(rng.standard_normal((B, n)))
    g = _cuda(rng.standard_normal((B, 2 * n * J)))
    Fx = A.ModwtFn.apply(x, J, filt)
    Ftg = torch.empty_like(x)
    W.check(W.lib().wh_modwt_fwd_adj(g.data_ptr(), Ftg.data_ptr(), n, B, J, filt, W.WH_F64, None))
    lhs, rhs = torch.sum(Fx * g).item(), torch.sum(x * Ftg).item()
    assert abs(lhs - rhs) <= 1e-12 * torch.linalg.norm(Fx).item() * torch.linalg.norm(g).item()
    Rg = A.ImodwtFn.apply(g, n, J, filt)
    Rtx = torch.empty_like(g)
    W.check(W.lib().wh_modwt_inv_adj(x.data_ptr(), Rtx.data_ptr(), n, B, J, filt, W.WH_F64, None))
    lhs, rhs = torch.sum(Rg * x).item(), torch.sum(g * Rtx).item()
    assert abs(lhs - rhs) <= 1e-12 * torch.linalg.norm(g).item() * torch.linalg.norm(x).item()


def _loss_and_grads(x, lam, wgt, J, filt, ttype, hard=W.SOFT):
    xv = x.clone().requires_grad_(True)
    lv = lam.clone().requires_grad_(True)
    y = A.smooth_levels(xv, lv, J, filt, ttype, hard)
    loss = torch.sum(y * wgt)
    loss.backward()
    return loss.item(), xv.grad.detach(), lv.grad.detach()


def _loss(x, lam, wgt, J, filt, ttype, hard=W.SOFT):
    with torch.no_grad():
        return torch.sum(A.smooth_levels(x, lam, J, filt, ttype, hard) * wgt).item()


@pytest.mark.parametrize("filt", FILTERS)
@pytest.mark.parametrize("ttype", [W.MODWT, W.DWT])
def test_finite_differences(filt, ttype):
    n = 256
    J = min(O.levels(n, 0, filt), 6)
    xs, _ = noisy_doppler(n, 3.0, seed=40 + filt)
    x = _cuda(xs)
    # per-level thresholds between coefficient magnitudes, away from kinks
    c = (D.modwt_fwd(x, J, filt) if ttype == W.MODWT else D.dwt_fwd(x, J, filt)).cpu().numpy()
    lam = np.zeros(J)
    for l in range(J):
        d = np.abs(c[2 * n * l + n:2 * n * l + 2 * n]) if ttype == W.MODWT else \
            np.abs(c[(1 << l)::(2 << l)])
        srt = np.sort(d)
        if len(srt) > 3:
            # midpoint of the widest gap in the middle half of |d|
            lo, hi = len(srt) // 4, 3 * len(srt) // 4
            k = lo + int(np.argmax(np.diff(srt[lo:hi + 1])))
            lam[l] = 0.5 * (srt[k] + srt[k + 1])
        else:
            lam[l] = 0.5 * (srt[0] + srt[-1]) if len(srt) > 1 else 0.5 * srt[0]
        assert np.min(np.abs(d - lam[l])) > 1e-4  # no kink within reach of h
    lam_t = _cuda(lam)
    wgt = _cuda(np.random.Generator(np.random.PCG64(7)).standard_normal(n))
    _, gx, glam = _loss_and_grads(x, lam_t, wgt, J, filt, ttype)
    h = 1e-6
    v = _cuda(np.random.Generator(np.random.PCG64(8)).standard_normal(n))
    fd = (_loss(x + h * v, lam_t, wgt, J, filt, ttype) -
          _loss(x - h * v, lam_t, wgt, J, filt, ttype)) / (2 * h)
    ad = torch.sum(gx * v).item()
    assert abs(fd - ad) <= 1e-6 * max(1.0, abs(ad)), (fd, ad)
    for l in range(J):
        e = torch.zeros(J, dtype=torch.float64, device="cuda")
        e[l] = h
        fdl = (_loss(x, lam_t + e, wgt, J, filt, ttype) -
               _loss(x, lam_t - e, wgt, J, filt, ttype)) / (2 * h)
        assert abs(fdl - glam[l].item()) <= 1e-6 * max(1.0, abs(fdl)), (l, fdl, glam[l].item())


def test_hard_threshold_gradient():
    n, J, filt = 1024, 6, 4
    x = _cuda(noisy_doppler(n, 2.0, seed=3)[0])
    lam = torch.full((J,), 0.05, dtype=torch.float64, device="cuda")
    w = D.modwt_fwd(x, J, filt)
    gy = _cuda(np.random.Generator(np.random.PCG64(2)).standard_normal(2 * n * J))
    gc = torch.empty_like(gy)
    glam = torch.full((J,), 7.0, dtype=torch.float64, device="cuda")
    thr = W.make_thresh(True, W.HARD, 0, J - 1, 0.0, lam.data_ptr())
    import ctypes as C
    W.check(W.lib().wh_thresh_grad(w.data_ptr(), gy.data_ptr(), gc.data_ptr(), glam.data_ptr(),
                                   n, 1, J, W.MODWT, filt, W.WH_F64, C.byref(thr), None))
    wc, gyc = w.cpu().numpy(), gy.cpu().numpy()
    ref = gyc.copy()
    for l in range(J):
        blk = slice(2 * n * l + n, 2 * n * l + 2 * n)
        ref[blk] = np.where(np.abs(wc[blk]) < 0.05, 0.0, gyc[blk])
    np.testing.assert_array_equal(gc.cpu().numpy(), ref)
    assert torch.count_nonzero(glam).item() == 0


def test_fit_level_thresholds_beats_universal():
    n, B, J, filt = 1 << 12, 16, 8, 4
    xs, clean = [], []
    for b in range(B):
        a, c = noisy_doppler(n, 3.0, seed=200 + b)
        xs.append(a)
        clean.append(c)
    x = _cuda(np.stack(xs), torch.float32)
    cl = _cuda(np.stack(clean), torch.float32)
    lam, losses = A.fit_level_thresholds(x, cl, J, filt, steps=60, lr=1e-2)
    print("loss", losses[0], "->", losses[-1], "lam", lam.cpu().numpy(), flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < losses[0]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("hardness", [W.SOFT, W.HARD])
def test_threshold_inverse_matches_composite(dtype, hardness):
    """The fused threshold + inverse (thresholds applied as the inverse reads)
    gives the composite threshold-then-inverse's output and gradients bit for
    bit, in the coefficients and in the thresholds."""
    B, n, J, f = 3, 1 << 12, 6, W.LA8
    g = torch.Generator(device="cuda").manual_seed(7)
    x = torch.randn(B, n, dtype=dtype, device="cuda", generator=g)
    w = A.modwt(x, J, f).detach()
    lam = torch.tensor([0.3, 0.25, 0.2, 0.2, 0.15, 0.1], dtype=torch.float64, device="cuda")
    gy = torch.randn(B, n, dtype=dtype, device="cuda", generator=g)
    outs = []
    for fused in (False, True):
        c = w.clone().requires_grad_(True)
        lv = lam.clone().requires_grad_(True)
        if fused:
            y = A.ThresholdInverseFn.apply(c, lv, n, J, f, hardness, 0, J - 1)
        else:
            y = A.ImodwtFn.apply(A.ThresholdFn.apply(c, lv, n, J, W.MODWT, f, hardness, 0, J - 1),
                                 n, J, f)
        y.backward(gy)
        outs.append((y.detach(), c.grad, lv.grad))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_threshold_grad_without_gc():
    """wh_thresh_grad with gc NULL gives the same per-level gradient."""
    n, J = 1 << 12, 5
    c = _cuda(white_noise(2 * n * J, seed=1))
    gy = _cuda(white_noise(2 * n * J, seed=2))
    lam = torch.full((J,), 0.5, dtype=torch.float64, device="cuda")
    thr = W.make_thresh(True, W.SOFT, 0, J - 1, 0.0, lam.data_ptr())
    gl = [torch.zeros(J, dtype=torch.float64, device="cuda") for _ in range(2)]
    gc = torch.empty_like(c)
    for k, gcp in enumerate((gc.data_ptr(), None)):
        W.check(W.lib().wh_thresh_grad(c.data_ptr(), gy.data_ptr(), gcp, gl[k].data_ptr(), n, 1,
                                       J, W.MODWT, W.LA8, W.WH_F64, C.byref(thr), None),
                "wh_thresh_grad")
    assert torch.equal(gl[0], gl[1])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("ttype", [W.MODWT, W.DWT])
def test_threshold_grad_many_chunks(ttype, dtype):
    """Per-level threshold gradients of several long signals (each level
    split over many workgroups) against a direct sum, and gc bit for bit."""
    n, J, count = 1 << 16, 6, 3
    per = 2 * n * J if ttype == W.MODWT else n
    rng = np.random.Generator(np.random.PCG64(31))
    c_np = rng.standard_normal(per * count).astype(np.float64 if dtype == torch.float64 else np.float32)
    gy_np = rng.standard_normal(per * count).astype(c_np.dtype)
    lam_np = np.linspace(0.3, 0.9, J)
    c, gy = torch.tensor(c_np, device="cuda"), torch.tensor(gy_np, device="cuda")
    lam = torch.tensor(lam_np, device="cuda")
    gc = torch.empty_like(c)
    glam = torch.zeros(J, dtype=torch.float64, device="cuda")
    thr = W.make_thresh(True, W.SOFT, 0, J - 1, 0.0, lam.data_ptr())
    W.check(W.lib().wh_thresh_grad(c.data_ptr(), gy.data_ptr(), gc.data_ptr(), glam.data_ptr(),
                                   n, count, J, ttype, W.LA8,
                                   W.WH_F64 if dtype == torch.float64 else W.WH_F32,
                                   C.byref(thr), None), "wh_thresh_grad")
    want_gc = gy_np.copy()
    want_gl = np.zeros(J)
    for s in range(count):
        base = s * per
        for l in range(J):
            if ttype == W.MODWT:
                idx = base + 2 * n * l + n + np.arange(n)
            else:
                idx = base + (1 << l) + (2 << l) * np.arange(n >> (l + 1))
            cv, g = c_np[idx].astype(np.float64), gy_np[idx].astype(np.float64)
            t = np.float64(c_np.dtype.type(lam_np[l]))
            keep = np.abs(cv) > t
            want_gc[idx] = np.where(keep, gy_np[idx], 0)
            want_gl[l] += np.sum(np.where(keep, np.where(cv < 0, g, -g), 0.0))
    np.testing.assert_array_equal(gc.cpu().numpy(), want_gc)
    np.testing.assert_allclose(glam.cpu().numpy(), want_gl, rtol=1e-12, atol=1e-9)
