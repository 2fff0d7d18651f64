"""Differentiable transforms and thresholding (torch.autograd over the C ABI).

Every backward is a hand-written adjoint kernel of libwavehip:

  modwt(x)            forward MODWT        backward: wh_modwt_fwd_adj
  imodwt(w, n)        inverse MODWT        backward: wh_modwt_inv_adj
  dwt(x), idwt(y)     DWT / inverse        backward: the orthogonal transpose
  threshold(c, lam)   per-level hard/soft  backward: wh_thresh_grad (c and lam)

x may be one signal (n,) or a batch (B, n) of equal-length signals; the
MODWT coefficients of a batch are (B, 2 n J) in the reference's
time-ordered layout.  `fit_level_thresholds` fits one soft threshold per
level by Adam, minimising the mean squared error against a clean signal.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as W
from .device import _dtype, _stream, filter_code, resolve_levels


def _rows(x: torch.Tensor):
    if x.dim() == 1:
        return x.unsqueeze(0), True
    if x.dim() == 2:
        return x, False
    raise W.WaveHipError("expected a signal (n,) or a batch (B, n)")


def _vp(t):
    return C.c_void_p(t.data_ptr())


class ModwtFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, J, filt):
        x2, squeeze = _rows(x.contiguous())
        B, n = x2.shape
        w = torch.empty(B, 2 * n * J, dtype=x.dtype, device=x.device)
        W.check(W.lib().wh_modwt_fwd_many(_vp(x2), _vp(w), n, B, J, filt, _dtype(x), _stream()),
                "wh_modwt_fwd_many")
        ctx.meta = (n, B, J, filt, squeeze)
        return w[0] if squeeze else w

    @staticmethod
    def backward(ctx, g):
        n, B, J, filt, squeeze = ctx.meta
        g2 = g.contiguous().view(B, 2 * n * J)
        gx = torch.empty(B, n, dtype=g.dtype, device=g.device)
        W.check(W.lib().wh_modwt_fwd_adj(_vp(g2), _vp(gx), n, B, J, filt, _dtype(g2), _stream()),
                "wh_modwt_fwd_adj")
        return (gx[0] if squeeze else gx), None, None


class ImodwtFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, n, J, filt):
        w2, squeeze = _rows(w.contiguous())
        B = w2.shape[0]
        if w2.shape[1] != 2 * n * J:
            raise W.WaveHipError("coefficient rows must hold 2 n J values")
        x = torch.empty(B, n, dtype=w.dtype, device=w.device)
        W.check(W.lib().wh_modwt_inv_many(_vp(w2), _vp(x), n, B, J, filt, _dtype(w), None,
                                          _stream()), "wh_modwt_inv_many")
        ctx.meta = (n, B, J, filt, squeeze)
        return x[0] if squeeze else x

    @staticmethod
    def backward(ctx, g):
        n, B, J, filt, squeeze = ctx.meta
        g2 = g.contiguous().view(B, n)
        gw = torch.empty(B, 2 * n * J, dtype=g.dtype, device=g.device)
        W.check(W.lib().wh_modwt_inv_adj(_vp(g2), _vp(gw), n, B, J, filt, _dtype(g2), _stream()),
                "wh_modwt_inv_adj")
        return (gw[0] if squeeze else gw), None, None, None


def _dwt_rows(fn, name, x, J, filt):
    x2, squeeze = _rows(x.contiguous())
    out = torch.empty_like(x2)
    for b in range(x2.shape[0]):
        W.check(fn(_vp(x2[b]), _vp(out[b]), x2.shape[1], J, filt, _dtype(x2), _stream()), name)
    return out[0] if squeeze else out


class DwtFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, J, filt):
        ctx.meta = (J, filt)
        return _dwt_rows(W.lib().wh_dwt_fwd, "wh_dwt_fwd", x, J, filt)

    @staticmethod
    def backward(ctx, g):
        J, filt = ctx.meta
        return _dwt_rows(W.lib().wh_dwt_fwd_adj, "wh_dwt_fwd_adj", g, J, filt), None, None


class IdwtFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, J, filt):
        ctx.meta = (J, filt)
        x2, squeeze = _rows(y.contiguous())
        out = torch.empty_like(x2)
        for b in range(x2.shape[0]):
            W.check(W.lib().wh_dwt_inv(_vp(x2[b]), _vp(out[b]), x2.shape[1], J, filt, _dtype(x2),
                                       None, _stream()), "wh_dwt_inv")
        return out[0] if squeeze else out

    @staticmethod
    def backward(ctx, g):
        J, filt = ctx.meta
        return _dwt_rows(W.lib().wh_dwt_inv_adj, "wh_dwt_inv_adj", g, J, filt), None, None


class ThresholdFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, lam, n, J, ttype, filt, hardness, minlevel, maxlevel):
        c2, squeeze = _rows(c.contiguous())
        if c2.shape[1] != (n if ttype == W.DWT else 2 * n * J):
            raise W.WaveHipError("coefficient rows must hold n (DWT) or 2 n J (MODWT) values")
        lam64 = lam.detach().to(device=c.device, dtype=torch.float64).contiguous()
        if lam64.numel() != J:
            raise W.WaveHipError("one threshold per level expected")
        y = c2.clone()
        thr = W.make_thresh(True, hardness, minlevel, maxlevel, 0.0, lam64.data_ptr())
        W.check(W.lib().wh_threshold_many(_vp(y), n, y.shape[0], J, ttype, filt, _dtype(y),
                                          C.byref(thr), _stream()), "wh_threshold_many")
        ctx.save_for_backward(c2, lam64)
        ctx.meta = (n, J, ttype, filt, hardness, minlevel, maxlevel, squeeze, lam.dtype)
        return y[0] if squeeze else y

    @staticmethod
    def backward(ctx, g):
        c2, lam64 = ctx.saved_tensors
        n, J, ttype, filt, hardness, minlevel, maxlevel, squeeze, lam_dtype = ctx.meta
        g2 = g.contiguous().view_as(c2)
        gc, glam = _thresh_grads(ctx, c2, g2, lam64, n, J, ttype, filt, hardness, minlevel,
                                 maxlevel)
        if gc is not None and squeeze:
            gc = gc[0]
        return (gc, None if glam is None else glam.to(lam_dtype), None, None, None, None, None,
                None, None)


def _thresh_grads(ctx, c2, g2, lam64, n, J, ttype, filt, hardness, minlevel, maxlevel):
    """wh_thresh_grad for the cotangents autograd asks for (ctx inputs 0 and
    1: the coefficients and the thresholds); None for the others."""
    want_c, want_lam = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
    if not (want_c or want_lam):
        return None, None
    gc = torch.empty_like(c2) if want_c else None
    glam = torch.zeros(J, dtype=torch.float64, device=c2.device) if want_lam else None
    thr = W.make_thresh(True, hardness, minlevel, maxlevel, 0.0, lam64.data_ptr())
    W.check(W.lib().wh_thresh_grad(_vp(c2), _vp(g2), None if gc is None else _vp(gc),
                                   None if glam is None else _vp(glam), n, c2.shape[0], J, ttype,
                                   filt, _dtype(c2), C.byref(thr), _stream()), "wh_thresh_grad")
    return gc, glam


class ThresholdInverseFn(torch.autograd.Function):
    """y = inverse MODWT of the coefficients with per-level thresholds applied
    as the inverse reads them (one pass; the same operations as thresholding
    first, so the same y bit for bit).  Backward: the inverse's adjoint, then
    the threshold's gradient."""

    @staticmethod
    def forward(ctx, c, lam, n, J, filt, hardness, minlevel, maxlevel):
        c2, squeeze = _rows(c.contiguous())
        B = c2.shape[0]
        if c2.shape[1] != 2 * n * J:
            raise W.WaveHipError("coefficient rows must hold 2 n J values")
        lam64 = lam.detach().to(device=c.device, dtype=torch.float64).contiguous()
        if lam64.numel() != J:
            raise W.WaveHipError("one threshold per level expected")
        thr = W.make_thresh(True, hardness, minlevel, maxlevel, 0.0, lam64.data_ptr())
        x = torch.empty(B, n, dtype=c.dtype, device=c.device)
        W.check(W.lib().wh_modwt_inv_many(_vp(c2), _vp(x), n, B, J, filt, _dtype(c), C.byref(thr),
                                          _stream()), "wh_modwt_inv_many")
        ctx.save_for_backward(c2, lam64)
        ctx.meta = (n, J, filt, hardness, minlevel, maxlevel, squeeze, lam.dtype)
        return x[0] if squeeze else x

    @staticmethod
    def backward(ctx, g):
        c2, lam64 = ctx.saved_tensors
        n, J, filt, hardness, minlevel, maxlevel, squeeze, lam_dtype = ctx.meta
        B = c2.shape[0]
        g2 = g.contiguous().view(B, n)
        gw = torch.empty(B, 2 * n * J, dtype=g.dtype, device=g.device)
        W.check(W.lib().wh_modwt_inv_adj(_vp(g2), _vp(gw), n, B, J, filt, _dtype(g2), _stream()),
                "wh_modwt_inv_adj")
        gc, glam = _thresh_grads(ctx, c2, gw, lam64, n, J, W.MODWT, filt, hardness, minlevel,
                                 maxlevel)
        if gc is not None and squeeze:
            gc = gc[0]
        return (gc, None if glam is None else glam.to(lam_dtype), None, None, None, None, None,
                None)


def modwt(x, nlevels, filter="LA8"):
    f = filter_code(filter)
    n = x.shape[-1]
    return ModwtFn.apply(x, resolve_levels(n, nlevels, f), f)


def imodwt(w, n, nlevels, filter="LA8"):
    f = filter_code(filter)
    return ImodwtFn.apply(w, n, resolve_levels(n, nlevels, f), f)


def dwt(x, nlevels, filter="LA8"):
    f = filter_code(filter)
    return DwtFn.apply(x, resolve_levels(x.shape[-1], nlevels, f), f)


def idwt(y, nlevels, filter="LA8"):
    f = filter_code(filter)
    return IdwtFn.apply(y, resolve_levels(y.shape[-1], nlevels, f), f)


def threshold(c, lam, n, nlevels, ttype, filter="LA8", hardness=W.SOFT, minlevel=0,
              maxlevel=None):
    """Threshold the details of levels [minlevel, maxlevel] of transformed
    signals c (DWT: (.., n); MODWT: (.., 2 n J)) by the per-level
    thresholds lam (J values); differentiable in c and lam."""
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    if maxlevel is None:
        maxlevel = J - 1
    return ThresholdFn.apply(c, lam, n, J, int(ttype), f, int(hardness), int(minlevel),
                             int(maxlevel))


def smooth_levels(x, lam, nlevels, filter="LA8", ttype=W.MODWT, hardness=W.SOFT):
    """Forward transform -> per-level threshold lam -> inverse, differentiable
    in x and lam."""
    f = filter_code(filter)
    n = x.shape[-1]
    J = resolve_levels(n, nlevels, f)
    if ttype == W.MODWT:
        w = ModwtFn.apply(x, J, f)
        return ThresholdInverseFn.apply(w, lam, n, J, f, hardness, 0, J - 1)
    y = DwtFn.apply(x, J, f)
    return IdwtFn.apply(ThresholdFn.apply(y, lam, n, J, W.DWT, f, hardness, 0, J - 1), J, f)


def fit_level_thresholds(x, clean, nlevels, filter="LA8", lam0=None, steps=300, lr=1e-2,
                         ttype=W.MODWT):
    """Fit one soft threshold per level by Adam on mean((smooth(x) - clean)^2).

    x, clean: (B, n) or (n,) on the GPU.  lam0: initial thresholds (J,),
    by default the universal threshold of the first signal.  Returns
    (lam, losses) with losses[i] the loss before step i and losses[-1] the
    loss after the last step."""
    from . import device as D

    f = filter_code(filter)
    n = x.shape[-1]
    J = resolve_levels(n, nlevels, f)
    if lam0 is None:
        x0 = x if x.dim() == 1 else x[0]
        w0 = D.modwt_fwd(x0.contiguous(), J, f) if ttype == W.MODWT else D.dwt_fwd(x0, J, f)
        t = D.univ_thresh(w0, n, J, ttype, f, 0, 0).item()
        lam0 = torch.full((J,), t, dtype=torch.float64)
    lam = lam0.detach().to(device=x.device, dtype=torch.float64).clone().requires_grad_(True)
    opt = torch.optim.Adam([lam], lr=lr)
    xd = x.detach()
    if ttype == W.MODWT:
        # the signals do not change: transform them once, and let each step
        # run the thresholding inverse and its adjoint only
        w = ModwtFn.apply(xd, J, f)

        def smooth(lv):
            return ThresholdInverseFn.apply(w, lv, n, J, f, W.SOFT, 0, J - 1)
    else:
        def smooth(lv):
            return smooth_levels(xd, lv, J, f, ttype)
    losses = []  # kept on the device: no host round trip per step
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = torch.mean((smooth(lam) - clean) ** 2)
        loss.backward()
        opt.step()
        losses.append(loss.detach())
    with torch.no_grad():
        losses.append(torch.mean((smooth(lam) - clean) ** 2))
    return lam.detach(), torch.stack(losses).double().cpu().tolist()
