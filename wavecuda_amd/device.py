"""Tensor-level API over the C ABI: torch provides device memory and the
stream, libwavehip.so does all the arithmetic.

Every function requires CUDA (HIP) tensors; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as W

_DT = {torch.float64: W.WH_F64, torch.float32: W.WH_F32}


def filter_code(filter) -> int:
    if isinstance(filter, str):
        try:
            return W.FILTER_CODES[filter]
        except KeyError:
            raise ValueError(f"unknown filter {filter!r}") from None
    return int(filter)


def _dtype(t: torch.Tensor) -> int:
    if t.dtype not in _DT:
        raise TypeError(f"unsupported dtype {t.dtype}; use float64 or float32")
    return _DT[t.dtype]


def _check_dev(*ts):
    for t in ts:
        if not t.is_cuda:
            raise ValueError("wavecuda_amd needs tensors on the GPU (device='cuda')")
        if not t.is_contiguous():
            raise ValueError("tensors must be contiguous")


def _need(t: torch.Tensor, count: int, dtype, what: str):
    """t must be a contiguous GPU tensor of `dtype` holding at least `count`
    values: the kernels read or write that many through its raw pointer."""
    _check_dev(t)
    if t.dtype != dtype:
        raise ValueError(f"{what}: dtype {t.dtype} does not match {dtype}")
    if t.numel() < count:
        raise ValueError(f"{what}: holds {t.numel()} values, needs {count}")


def _coeff_count(ttype: int, n: int, J: int) -> int:
    """Values of a transformed vector: n (DWT) or 2 n J (MODWT)."""
    return n if int(ttype) == 0 else 2 * n * J


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def resolve_levels(n: int, nlevels: int, filter) -> int:
    J = W.lib().wh_resolve_levels(int(n), int(nlevels), filter_code(filter))
    if J <= 0:
        raise ValueError(f"invalid length {n} / nlevels {nlevels} for filter {filter}")
    return J


def _thresh_struct(thresh, J=None):
    if thresh is None:
        return None, None
    if isinstance(thresh, W.wh_thresh):
        return thresh, None
    value, hardness, minlevel, maxlevel, levels = (
        thresh.get("value", 0.0),
        thresh.get("hardness", W.SOFT),
        thresh.get("minlevel", 0),
        thresh.get("maxlevel", 0),
        thresh.get("levels"),
    )
    keep = None
    ptr = None
    if levels is not None:
        keep = levels.to(dtype=torch.float64).contiguous()
        _check_dev(keep)
        if J is not None and keep.numel() < J:
            raise ValueError(f"per-level thresholds: {keep.numel()} values for {J} levels")
        ptr = keep.data_ptr()
    t = W.make_thresh(True, hardness, minlevel, maxlevel, value, ptr)
    return t, keep


def modwt_fwd(x: torch.Tensor, nlevels: int = 0, filter="LA8", out: torch.Tensor | None = None):
    """Time-ordered MODWT: returns xmod of 2*n*J values [S1|D1|...|SJ|DJ]."""
    _check_dev(x)
    n = x.numel()
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    if out is None:
        out = torch.empty(2 * n * J, dtype=x.dtype, device=x.device)
    _need(out, 2 * n * J, x.dtype, "modwt_fwd out")
    W.check(W.lib().wh_modwt_fwd(x.data_ptr(), out.data_ptr(), n, J, f, _dtype(x), _stream()),
            "wh_modwt_fwd")
    return out


def modwt_inv(xmod: torch.Tensor, n: int, nlevels: int = 0, filter="LA8", thresh=None,
              out: torch.Tensor | None = None):
    """Inverse MODWT of xmod (2*n*J values) with optional fused thresholding.

    ``thresh`` is None or a dict(value=, hardness=, minlevel=, maxlevel=,
    levels=<device tensor of per-level thresholds>)."""
    _check_dev(xmod)
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    if xmod.numel() < 2 * n * J:
        raise ValueError("xmod is shorter than 2*n*nlevels")
    if out is None:
        out = torch.empty(n, dtype=xmod.dtype, device=xmod.device)
    _need(out, n, xmod.dtype, "modwt_inv out")
    t, keep = _thresh_struct(thresh, J)
    W.check(
        W.lib().wh_modwt_inv(xmod.data_ptr(), out.data_ptr(), n, J, f, _dtype(xmod),
                             C.byref(t) if t is not None else None, _stream()),
        "wh_modwt_inv")
    del keep
    return out


def dwt_fwd(x: torch.Tensor, nlevels: int = 0, filter="LA8", out: torch.Tensor | None = None):
    """Periodic DWT in the interleaved layout (out may be x for in place)."""
    _check_dev(x)
    n = x.numel()
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    if out is None:
        out = torch.empty_like(x)
    _need(out, n, x.dtype, "dwt_fwd out")
    W.check(W.lib().wh_dwt_fwd(x.data_ptr(), out.data_ptr(), n, J, f, _dtype(x), _stream()),
            "wh_dwt_fwd")
    return out


def dwt_inv(y: torch.Tensor, nlevels: int = 0, filter="LA8", thresh=None,
            out: torch.Tensor | None = None):
    """Inverse periodic DWT with optional fused thresholding."""
    _check_dev(y)
    n = y.numel()
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    if out is None:
        out = torch.empty_like(y)
    _need(out, n, y.dtype, "dwt_inv out")
    t, keep = _thresh_struct(thresh, J)
    W.check(
        W.lib().wh_dwt_inv(y.data_ptr(), out.data_ptr(), n, J, f, _dtype(y),
                           C.byref(t) if t is not None else None, _stream()),
        "wh_dwt_inv")
    del keep
    return out


def modwt_tune(n: int, nlevels: int, filter="LA8", dtype=torch.float64) -> None:
    """Time the fused MODWT's plan alternatives for this shape and keep the
    default plan for later calls unless another is more than 2% faster
    (synchronous; results do not depend on the plan)."""
    f = filter_code(filter)
    W.check(W.lib().wh_modwt_tune(int(n), resolve_levels(n, nlevels, f), f, _DT[dtype],
                                  _stream()), "wh_modwt_tune")


def modwt_plan(inverse: bool, n: int, nlevels: int, filter="LA8", dtype=torch.float64,
               count: int = 1) -> str:
    """The level groups the fused MODWT runs for this shape (a launch of
    `count` signals), as text."""
    buf = C.create_string_buffer(4096)
    rc = W.lib().wh_debug_modwt_plan_many(1 if inverse else 0, n, count, nlevels,
                                         filter_code(filter), _DT[dtype], buf, 4096)
    W.check(rc, "wh_debug_modwt_plan_many")
    return buf.value.decode()


def modwt_set_plan(inverse: bool, n: int, nlevels: int, filter="LA8", dtype=torch.float64,
                   knobs=None) -> None:
    """Set the planner's choices for one shape as the tuner would: knobs =
    (deepest first group, segments per CU of the first group, of deeper
    groups, threads of the first inverse group); None restores the default
    rules."""
    g1, f, d, nt = knobs if knobs is not None else (0, 1, 1, 512)
    W.check(W.lib().wh_debug_modwt_set_plan(1 if inverse else 0, n, nlevels, filter_code(filter),
                                            _DT[dtype], g1, f, d, nt), "wh_debug_modwt_set_plan")


def modwt_plan_candidates(inverse: bool, n: int, nlevels: int, filter="LA8",
                          dtype=torch.float64):
    """Every distinct plan the tuner times for this shape: {plan text: knobs}
    (the tuner's search space: deepest first group min(J, 5)..10, segments
    per CU of the first group 1/2/4 and of deeper groups 1/2/4/8, and for
    the inverse a 256-thread first group).  Leaves the default rules in
    force."""
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    out = {}
    try:
        for nt in ((512, 256) if inverse else (512,)):
            for g1 in range(min(J, 5), 11):
                for fp in (1, 2, 4):
                    for dp in (1, 2, 4, 8):
                        k = (g1, fp, dp, nt)
                        modwt_set_plan(inverse, n, J, f, dtype, k)
                        out.setdefault(modwt_plan(inverse, n, J, f, dtype), k)
    finally:
        modwt_set_plan(inverse, n, J, f, dtype, None)
    return out


def modwt_convert(xmod: torch.Tensor, n: int, nlevels: int, filter="LA8", to_po: bool = True,
                  out: torch.Tensor | None = None):
    """A MODWT coefficient vector (2nJ values) from time order to packet
    order (to_po) or back (convert_modwt, src/wavutils.cpp:532-600)."""
    _check_dev(xmod)
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    if xmod.numel() < 2 * n * J:
        raise ValueError(f"xmod must hold 2 n J = {2 * n * J} values")
    if out is None:
        out = torch.empty(2 * n * J, dtype=xmod.dtype, device=xmod.device)
    _need(out, 2 * n * J, xmod.dtype, "modwt_convert out")
    W.check(W.lib().wh_modwt_convert(xmod.data_ptr(), out.data_ptr(), n, J, f,
                                     1 if to_po else 0, _dtype(xmod), _stream()),
            "wh_modwt_convert")
    return out


def modwt_fwd_po(x: torch.Tensor, nlevels: int = 0, filter="LA8"):
    """Packet-ordered MODWT forward (src/transform.cpp:39-55)."""
    return modwt_convert(modwt_fwd(x, nlevels, filter), x.numel(), nlevels, filter, True)


def modwt_inv_po(xmod_po: torch.Tensor, n: int, nlevels: int = 0, filter="LA8", thresh=None):
    """Inverse of a packet-ordered MODWT vector."""
    return modwt_inv(modwt_convert(xmod_po, n, nlevels, filter, False), n, nlevels, filter,
                     thresh)


def coeff_level(coeffs: torch.Tensor, n: int, nlevels: int, ttype: int, filter, level: int,
                coeff_type: str = "d") -> torch.Tensor:
    """One level's coefficients of a transformed vector (DWT n / MODWT 2nJ
    values) as a new compact tensor in translate order: the values at the
    positions getCoeffLevel (R/rplotting.R:10-46) lists for a DWT, the
    level's D or S block for a time-ordered MODWT.  DWT scaling coefficients
    ("s") exist at the deepest level only."""
    _check_dev(coeffs)
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    if coeff_type not in ("d", "s"):
        raise ValueError('coeff_type must be "d" or "s"')
    _need(coeffs, _coeff_count(ttype, n, J), coeffs.dtype, "coeff_level coeffs")
    if not 1 <= level <= J:
        raise ValueError(f"level must be between 1 and {J}")
    s = coeff_type == "s"
    count = (n >> (J if s else level)) if int(ttype) == 0 else n
    out = torch.empty(max(count, 1), dtype=coeffs.dtype, device=coeffs.device)
    got = C.c_int64(0)
    W.check(W.lib().wh_coeff_level(coeffs.data_ptr(), out.data_ptr(), n, J, int(ttype), f,
                                   int(level), 1 if s else 0, _dtype(coeffs), C.byref(got),
                                   _stream()), "wh_coeff_level")
    return out[: got.value]


def coeff_labels(n: int, nlevels: int, ttype: int, filter="Haar"):
    """WSTtoDT's label columns (R/rfunctions.R:925-1004) for every
    coefficient in storage order, computed on the device: (level int32,
    translate float64, is_scaling uint8) tensors of length n (DWT) or 2nJ
    (MODWT)."""
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    m = n if int(ttype) == 0 else 2 * n * J
    level = torch.empty(m, dtype=torch.int32, device="cuda")
    translate = torch.empty(m, dtype=torch.float64, device="cuda")
    kind = torch.empty(m, dtype=torch.uint8, device="cuda")
    W.check(W.lib().wh_coeff_labels(n, J, int(ttype), f, level.data_ptr(), translate.data_ptr(),
                                    kind.data_ptr(), _stream()), "wh_coeff_labels")
    return level, translate, kind


def threshold_(coeffs: torch.Tensor, n: int, nlevels: int, ttype: int, filter, value: float,
               hardness: int, minlevel: int, maxlevel: int, levels: torch.Tensor | None = None):
    """In-place thresholding of a transformed vector (DWT n / MODWT 2nJ values)."""
    _check_dev(coeffs)
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    _need(coeffs, _coeff_count(ttype, n, J), coeffs.dtype, "threshold_ coeffs")
    t, keep = _thresh_struct(dict(value=value, hardness=hardness, minlevel=minlevel,
                                  maxlevel=maxlevel, levels=levels), J)
    W.check(W.lib().wh_threshold(coeffs.data_ptr(), n, J, int(ttype), f, _dtype(coeffs),
                                 C.byref(t), _stream()), "wh_threshold")
    del keep
    return coeffs


def _scalar_out(fn, what, *args):
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    W.check(fn(*args, out.data_ptr(), _stream()), what)
    return out


def median(x: torch.Tensor) -> torch.Tensor:
    """Exact median (average of the two middle order statistics), 1 double."""
    _check_dev(x)
    return _scalar_out(W.lib().wh_median, "wh_median", x.data_ptr(), x.numel(), _dtype(x))


def mad(x: torch.Tensor) -> torch.Tensor:
    """1.4826 * median(|x - median(x)|), 1 double."""
    _check_dev(x)
    return _scalar_out(W.lib().wh_mad, "wh_mad", x.data_ptr(), x.numel(), _dtype(x))


def univ_thresh(coeffs: torch.Tensor, n: int, nlevels: int, ttype: int, filter,
                minlevel: int = 0, maxlevel: int = 0) -> torch.Tensor:
    _check_dev(coeffs)
    f = filter_code(filter)
    J = resolve_levels(n, nlevels, f)
    _need(coeffs, _coeff_count(ttype, n, J), coeffs.dtype, "univ_thresh coeffs")
    return _scalar_out(W.lib().wh_univ_thresh, "wh_univ_thresh", coeffs.data_ptr(), n, J,
                       int(ttype), f, _dtype(coeffs), int(minlevel), int(maxlevel))


def interp_mse(xn: torch.Tensor, ye: torch.Tensor, yo: torch.Tensor) -> torch.Tensor:
    _check_dev(xn, ye, yo)
    h = xn.numel() // 2
    for t, name in ((ye, "ye"), (yo, "yo")):
        _need(t, h, xn.dtype, f"interp_mse {name}")
        if t.numel() != h:
            raise ValueError(f"interp_mse {name}: holds {t.numel()} values, needs n/2 = {h}")
    return _scalar_out(W.lib().wh_interp_mse, "wh_interp_mse", xn.data_ptr(), ye.data_ptr(),
                       yo.data_ptr(), xn.numel(), _dtype(xn))


def smooth(x: torch.Tensor, nlevels: int, ttype: int, filter, threshtype: int,
           thresh: float | None = None, hardness: int = 1, minlevel: int = 0,
           maxlevel: int | None = None, tol: float = 0.01, inplace: bool = False):
    """Forward -> threshold (0 manual / 1 universal / 2 CV) -> inverse.

    Returns (smoothed tensor, threshold used)."""
    _check_dev(x)
    f = filter_code(filter)
    n = x.numel()
    J = resolve_levels(n, nlevels, f)
    if maxlevel is None:
        maxlevel = J - 1
    y = x if inplace else x.clone()
    t = C.c_double(0.0 if thresh is None else float(thresh))
    W.check(W.lib().wh_smooth(y.data_ptr(), n, J, int(ttype), f, _dtype(y), int(threshtype),
                              C.byref(t), int(hardness), int(minlevel), int(maxlevel),
                              float(tol), _stream()), "wh_smooth")
    return y, t.value
