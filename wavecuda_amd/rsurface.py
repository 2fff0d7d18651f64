"""The reference package's R surface, restated in Python over the `.C` symbols.

GPUTransform, GPUReconstruct, GPUSmooth (R/rfunctions.R:383-406, :452-468,
:588-612), CPUTransform, CPUReconstruct, CPUThreshold, CPUSmooth (:334-357,
:421-437, :508-526, :547-567) and a GPU counterpart of CPUThreshold, with the same
argument names (R's dots become underscores), the same checks and error
messages (check.trans.inputs :64-119, check.thresh.inputs :168-202,
check.smooth.inputs :211-249) and the same returned WST object (return.trans
:259-292, return.thresh :300-308): fields x, ttype, filt, filtlen, nlevels,
len, w, where a MODWT's w is the 2 n log2(n) vector R allocates
(:94) with the 2 n J coefficients first.

The transforms and smoothing run through the `.C` symbols each R function
binds (RgpuTransform / RgpuSmooth for the GPU functions, RcpuTransform /
RcpuThreshold / RcpuSmooth for the CPU-named ones, which this library also
runs on the GPU) called exactly the way R's `.C` calls them (host double vectors, pointers
to one int each, results written in place), so this module also exercises
the drop-in boundary.  Differences from the reference, all additions: C6
and LA8 MODWT are accepted (the reference stops at :114-116), and
GPUSmooth accepts "univ" and every filter (the reference stops at
:592-594).  R's stop() is raised as ValueError.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as W

_TTYPE = {"DWT": 0, "MODWT": 1}
_FILT = {"Haar": 1, "D4": 2, "C6": 3, "LA8": 4}
_FLEN = {"Haar": 2, "D4": 4, "C6": 6, "LA8": 8}
_SENSE = {"FWD": 1, "BWD": 0}
_HARD = {"hard": 0, "soft": 1}
_THRESH = {"manual": 0, "univ": 1, "cv": 2}


@dataclass
class WST:
    """R's WST object (R/rfunctions.R:267-284)."""
    x: np.ndarray | None
    ttype: str
    filt: str
    filtlen: int
    nlevels: int
    len: int
    w: np.ndarray = field(repr=False)


def _is_pow2(v: float) -> bool:
    return v > 0 and math.log2(v) == round(math.log2(v))


def check_trans_inputs(xin, direction, nlevels, transform_type, filter):
    """check.trans.inputs (R/rfunctions.R:64-119)."""
    if transform_type not in _TTYPE:
        raise ValueError("Unexpected transform type")
    ttype = _TTYPE[transform_type]
    if filter not in _FILT:
        raise ValueError("Unexpected filter")
    filt, filtlen = _FILT[filter], _FLEN[filter]
    if direction not in _SENSE:
        raise ValueError("Unexpected direction!")
    sense = _SENSE[direction]
    xin = np.asarray(xin)
    if not np.issubdtype(xin.dtype, np.number):
        raise ValueError("Unexpected input vector")
    lenxin = xin.size
    if ttype == 0:
        x = np.array(xin, dtype=np.float64)
        xmod = np.zeros(0)
        length = lenxin
        if not _is_pow2(length):
            raise ValueError("Need length to be power of 2")
    elif sense == 1:
        x = np.array(xin, dtype=np.float64)
        length = lenxin
        if not _is_pow2(length):
            raise ValueError("Need length to be power of 2")
        xmod = np.zeros(int(length * 2 * math.log2(length)))
    else:
        xmod = np.array(xin, dtype=np.float64)
        if nlevels == 0:
            raise ValueError("Probably got an error of infinite length: len = Inf")
        length = lenxin / (2 * nlevels)
        if not _is_pow2(length):
            raise ValueError("Was this really a MODWT vector? If so, check nlevels. I calculate "
                             "the reconstruction vector to be of inappropriate size.")
        length = int(length)
        x = np.zeros(length)
    maxlevels = math.floor(math.log2(length / filtlen)) + 1
    if nlevels == 0:
        nlevels = maxlevels
    if nlevels != round(nlevels) or nlevels < 0 or nlevels > maxlevels:
        raise ValueError(f"nlevels should be an integer [whole number] between 1 and {maxlevels}")
    return {"x": x, "xmod": xmod, "len": int(length), "sense": sense, "nlevels": int(nlevels),
            "ttype": ttype, "filt": filt, "filtlen": filtlen}


def prepare_wst_arg_list(xwav: WST):
    """prepare.wst.arg.list (R/rfunctions.R:128-155)."""
    base = {"len": xwav.len, "sense": 0, "nlevels": xwav.nlevels,
            "ttype": _TTYPE[xwav.ttype], "filt": _FILT[xwav.filt], "filtlen": xwav.filtlen}
    if xwav.ttype == "DWT":
        return {"x": np.array(xwav.w, dtype=np.float64), "xmod": np.zeros(0), **base}
    x = np.array(xwav.x if xwav.x is not None else np.zeros(xwav.len), dtype=np.float64)
    return {"x": x, "xmod": np.array(xwav.w, dtype=np.float64), **base}


def check_thresh_inputs(xwav: WST, hard_soft, thresh, min_level, max_level):
    """check.thresh.inputs (R/rfunctions.R:168-202)."""
    a = prepare_wst_arg_list(xwav)
    if np.ndim(thresh) != 0:
        raise ValueError("Only scalar thresholds supported")
    if thresh <= 0:
        raise ValueError("Threshold must be greater than 0")
    a["thresh"] = float(thresh)
    del a["sense"]
    if hard_soft not in _HARD:
        raise ValueError("Only hard or soft thresholding supported")
    a["hardness"] = _HARD[hard_soft]
    if (min_level != round(min_level) or max_level != round(max_level) or min_level > max_level
            or min_level <= 0 or max_level > xwav.nlevels):
        raise ValueError("min.level and max.level must be integers in the range 1...nlevels")
    a["min_level"] = int(min_level) - 1
    a["max_level"] = int(max_level) - 1
    return a


def check_smooth_inputs(xin, nlevels, transform_type, filter, thresh_type, thresh, hard_soft,
                        min_level, max_level, tol):
    """check.smooth.inputs (R/rfunctions.R:211-249)."""
    a = check_trans_inputs(xin, "FWD", nlevels, transform_type, filter)
    del a["sense"]
    if thresh_type not in _THRESH:
        raise ValueError("Only manual, univ or cv supported for threshold type")
    a["threshtype"] = _THRESH[thresh_type]
    if thresh_type == "manual":
        if thresh is None or np.ndim(thresh) != 0:
            raise ValueError("Only scalar thresholds supported")
        if thresh <= 0:
            raise ValueError("Threshold must be greater than 0")
        a["thresh"] = float(thresh)
    else:
        a["thresh"] = None
    if thresh_type == "cv":
        if tol <= 0:
            raise ValueError("Tolerance must be greater than 0")
        a["tol"] = float(tol)
    else:
        a["tol"] = None
    if hard_soft not in _HARD:
        raise ValueError("Only hard or soft thresholding supported")
    a["hardness"] = _HARD[hard_soft]
    if (min_level != round(min_level) or max_level != round(max_level) or min_level > max_level
            or min_level <= 0 or max_level > nlevels):
        raise ValueError("min.level and max.level must be integers in the range 1...nlevels")
    a["min_level"] = int(min_level) - 1
    a["max_level"] = int(max_level) - 1
    return a


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(v):
    return C.byref(C.c_int(int(v)))


def _call_transform(a, sym="RgpuTransform"):
    """.C(sym, x, xmod, len, sense, nlevels, ttype, filter, filterlen) --
    R/rfunctions.R:345-354 (RcpuTransform), :394-403 (RgpuTransform)."""
    x = np.ascontiguousarray(a["x"], dtype=np.float64)
    xmod = np.ascontiguousarray(a["xmod"], dtype=np.float64)
    getattr(W.lib(), sym)(_dp(x), _dp(xmod) if xmod.size else None, _ip(a["len"]),
                          _ip(a["sense"]), _ip(a["nlevels"]), _ip(a["ttype"]), _ip(a["filt"]),
                          _ip(a["filtlen"]))
    out = dict(a)
    out["x"], out["xmod"] = x, xmod
    return out


def return_trans(arglist, argsin):
    """return.trans (R/rfunctions.R:259-292)."""
    if arglist["sense"] == 1:
        w = arglist["x"] if argsin["ttype"] == "DWT" else arglist["xmod"]
        return WST(x=argsin.get("x"), ttype=argsin["ttype"], filt=argsin["filter"],
                   filtlen=arglist["filtlen"], nlevels=arglist["nlevels"], len=arglist["len"],
                   w=w)
    return arglist["x"]


def GPUTransform(xin, direction="FWD", nlevels=0, transform_type="DWT", filter="Haar"):
    """GPUTransform (R/rfunctions.R:383-406): a WST for FWD, a vector for BWD."""
    args_in = {"ttype": transform_type, "filter": filter}
    if direction == "FWD":
        args_in["x"] = np.asarray(xin)
    a = check_trans_inputs(xin, direction, nlevels, transform_type, filter)
    return return_trans(_call_transform(a), args_in)


def GPUReconstruct(xwav: WST):
    """GPUReconstruct (R/rfunctions.R:452-468)."""
    return _call_transform(prepare_wst_arg_list(xwav))["x"]


def CPUTransform(xin, direction="FWD", nlevels=0, transform_type="DWT", filter="Haar"):
    """CPUTransform (R/rfunctions.R:334-357) over RcpuTransform."""
    args_in = {"ttype": transform_type, "filter": filter}
    if direction == "FWD":
        args_in["x"] = np.asarray(xin)
    a = check_trans_inputs(xin, direction, nlevels, transform_type, filter)
    return return_trans(_call_transform(a, "RcpuTransform"), args_in)


def CPUReconstruct(xwav: WST):
    """CPUReconstruct (R/rfunctions.R:421-437) over RcpuTransform."""
    return _call_transform(prepare_wst_arg_list(xwav), "RcpuTransform")["x"]


def return_thresh(xwav: WST, arg_list):
    """return.thresh (R/rfunctions.R:300-308)."""
    out = WST(**{**xwav.__dict__})
    out.w = arg_list["x"] if xwav.ttype == "DWT" else arg_list["xmod"]
    return out


def CPUThreshold(xwav: WST, thresh, hard_soft, min_level, max_level):
    """CPUThreshold (R/rfunctions.R:508-526) over RcpuThreshold."""
    a = check_thresh_inputs(xwav, hard_soft, thresh, min_level, max_level)
    x = np.ascontiguousarray(a["x"], dtype=np.float64)
    xmod = np.ascontiguousarray(a["xmod"], dtype=np.float64)
    t = np.array([a["thresh"]])
    W.lib().RcpuThreshold(_dp(x), _dp(xmod) if xmod.size else None, _ip(xwav.len),
                          _ip(xwav.nlevels), _ip(a["ttype"]), _ip(a["filt"]), _ip(a["filtlen"]),
                          _dp(t), _ip(a["hardness"]), _ip(a["min_level"]), _ip(a["max_level"]))
    return return_thresh(xwav, {"x": x, "xmod": xmod})


def GPUThreshold(xwav: WST, thresh, hard_soft, min_level, max_level):
    """The GPU counterpart of CPUThreshold (R/rfunctions.R:508-526): the
    WST with its coefficients thresholded on the device."""
    import torch

    from . import device as D

    a = check_thresh_inputs(xwav, hard_soft, thresh, min_level, max_level)
    w = torch.tensor(np.asarray(xwav.w, dtype=np.float64), device="cuda")
    D.threshold_(w, xwav.len, xwav.nlevels, a["ttype"], a["filt"], a["thresh"], a["hardness"],
                 a["min_level"], a["max_level"])
    out = WST(**{**xwav.__dict__})
    out.w = w.cpu().numpy()
    return out


def CPUSmooth(xin, nlevels, transform_type, filter, thresh_type, thresh=None, hard_soft="soft",
              min_level=1, max_level=None, tol=0.01):
    """CPUSmooth (R/rfunctions.R:547-567) over RcpuSmooth: the smoothed vector."""
    return GPUSmooth(xin, nlevels, transform_type, filter, thresh_type, thresh, hard_soft,
                     min_level, max_level, tol, _sym="RcpuSmooth")


def GPUSmooth(xin, nlevels, transform_type, filter, thresh_type, thresh=None, hard_soft="soft",
              min_level=1, max_level=None, tol=0.01, _sym="RgpuSmooth"):
    """GPUSmooth (R/rfunctions.R:588-612): the smoothed vector."""
    if max_level is None:
        max_level = nlevels
    a = check_smooth_inputs(xin, nlevels, transform_type, filter, thresh_type, thresh, hard_soft,
                            min_level, max_level, tol)
    x = np.ascontiguousarray(a["x"], dtype=np.float64)
    t = np.array([a["thresh"] if a["thresh"] is not None else 0.0])
    tl = np.array([a["tol"] if a["tol"] is not None else 0.01])
    getattr(W.lib(), _sym)(_dp(x), _ip(a["len"]), _ip(a["nlevels"]), _ip(a["ttype"]),
                           _ip(a["filt"]), _ip(a["filtlen"]), _ip(a["threshtype"]), _dp(t),
                           _ip(a["hardness"]), _ip(a["min_level"]), _ip(a["max_level"]), _dp(tl))
    return x


# ---- coefficients by level: getCoeffLevel, label.detail.scaling, WSTtoDT --

def getCoeffLevel(xw: WST, level, coeffType="d"):
    """getCoeffLevel (R/rplotting.R:10-46): the 1-based positions of level
    `level`'s detail ("d") or scaling ("s") coefficients in xw.w.  As in the
    R code, a MODWT gives 1..len whatever the level and type; the device
    gather `device.coeff_level` returns the level's own D or S block."""
    if not isinstance(xw, WST):
        raise ValueError("WST type required")
    if level <= 0 or level > xw.nlevels or level != int(level):
        raise ValueError(f"Integer level expected, between 1 and {xw.nlevels}")
    if coeffType not in ("d", "s"):
        raise ValueError('coeffType must be "d" or "s"')
    d = coeffType == "d"
    n = xw.len
    if xw.ttype == "DWT":
        if not d and level != xw.nlevels:
            raise ValueError("scaling coefficients exist at level nlevels only")
        start, gap = 2 ** (int(level) - 1) * d + 1, 2 ** int(level)
    else:
        start, gap = 1, 1
    return np.arange(start, n + 1, gap, dtype=np.int64)


def _labels(xwav: WST):
    """(level, translate, is_scaling) of every coefficient, from the device."""
    from . import device as D

    lev, tr, kind = D.coeff_labels(xwav.len, xwav.nlevels, _TTYPE[xwav.ttype], xwav.filt)
    return lev.cpu().numpy(), tr.cpu().numpy(), kind.cpu().numpy()


def label_detail_scaling(Xwav: WST):
    """label.detail.scaling (R/rfunctions.R:902-912): "s" or "d" for each
    coefficient in storage order (len of them for a DWT, 2 len nlevels for a
    MODWT), computed on the device."""
    kind = _labels(Xwav)[2]
    return np.where(kind == 1, "s", "d")


def WSTtoDT(Xwav: WST, scaling=True, forPlotting=False):
    """WSTtoDT (R/rfunctions.R:925-1004) as a pandas DataFrame with the
    columns W, CoeffType, Level, Translate (and minmax when forPlotting),
    one row per coefficient in storage order; the label columns come from
    the device.  forPlotting drops the scaling rows and appends, for every
    level, two rows of Translate 0 holding -max|W| and +max|W| of that
    level's details (minmax 1 and 2).  A MODWT's rows are its 2 len nlevels
    coefficients (R pads w to 2 len log2(len))."""
    import pandas as pd

    lev, tr, kind = _labels(Xwav)
    w = np.asarray(Xwav.w, dtype=np.float64)[: lev.size]
    if forPlotting:
        scaling = False
    df = pd.DataFrame({"W": w, "CoeffType": np.where(kind == 1, "s", "d"),
                       "Level": lev.astype(np.int64), "Translate": tr,
                       "minmax": np.zeros(lev.size, dtype=np.int64)})
    if forPlotting:
        pads = []
        for level in range(1, Xwav.nlevels + 1):
            sel = (df["CoeffType"] == "d") & (df["Level"] == level)
            m = float(np.abs(df.loc[sel, "W"]).max()) if sel.any() else np.nan
            for mm, sign in ((1, -1.0), (2, 1.0)):
                pads.append({"W": sign * m, "CoeffType": "d", "Level": level,
                             "Translate": 0.0, "minmax": mm})
        df = pd.concat([df, pd.DataFrame(pads)], ignore_index=True)
    else:
        df = df.drop(columns="minmax")
    if not scaling:
        df = df[df["CoeffType"] == "d"].reset_index(drop=True)
    return df


# ---- wavethresh interop: WSTtoWavethresh -----------------------------------

def wavethresh_layout(n, nlevels, ttype, filt):
    """The calls WSTtoWavethresh (R/rfunctions.R:864-893) makes on a zero
    wavethresh wd object, from wh_wavethresh_layout: a dict with the wd
    arguments (filter.number, family, type, bc), the warning flags, and the
    putD / putC calls in order as (wavethresh level, "d" or "s", offset,
    count) into the packed vector of wh_wavethresh_pack."""
    cap = 64
    table = (W.wh_wt_entry * cap)()
    fnum = C.c_int(0)
    fam = C.create_string_buffer(32)
    warn = C.c_int(0)
    k = W.lib().wh_wavethresh_layout(int(n), int(nlevels), _TTYPE[ttype], _FILT[filt], table,
                                     cap, C.byref(fnum), fam, 32, C.byref(warn))
    if k <= 0:
        raise ValueError(W.lib().wh_last_error().decode())
    calls = [(table[i].wt_level, "s" if table[i].coeff_type else "d", table[i].offset,
              table[i].count) for i in range(k)]
    return {"filter.number": fnum.value, "family": fam.value.decode(),
            "type": "wavelet" if ttype == "DWT" else "station", "bc": "periodic",
            "warnings": warn.value, "calls": calls}


def WSTtoWavethresh(XW: WST, showWarnings=True):
    """WSTtoWavethresh (R/rfunctions.R:864-893): what it hands to wavethresh
    -- the wd() arguments and, in call order, each putD(level, v) and the
    final putC(level, v) -- with v gathered on the GPU by
    wh_wavethresh_pack.  (The wd object itself is built by the wavethresh R
    package, which is not part of this library.)  The reference's warnings
    are raised as UserWarning."""
    import warnings

    import torch

    lay = wavethresh_layout(XW.len, XW.nlevels, XW.ttype, XW.filt)
    if showWarnings and (lay["warnings"] & 1):
        warnings.warn("Coiflets not implemented in wavethresh; you will not be able reconstruct "
                      "with wavethresh")
    if showWarnings and (lay["warnings"] & 2):
        warnings.warn("Wavethresh uses a different ordering in MODWT coefficients")
    n, J = XW.len, XW.nlevels
    m = 2 * n * J if XW.ttype == "MODWT" else n
    w = torch.tensor(np.asarray(XW.w[:m], dtype=np.float64), device="cuda")
    total = sum(c[3] for c in lay["calls"])
    out = torch.empty(total, dtype=torch.float64, device="cuda")
    W.check(W.lib().wh_wavethresh_pack(w.data_ptr(), out.data_ptr(), n, J, _TTYPE[XW.ttype],
                                       _FILT[XW.filt], W.WH_F64, None), "wh_wavethresh_pack")
    v = out.cpu().numpy()
    puts = [{"fn": "putD" if t == "d" else "putC", "level": lev, "v": v[o:o + c]}
            for lev, t, o, c in lay["calls"]]
    return {"wd": {"n": n, "filter.number": lay["filter.number"], "family": lay["family"],
                   "type": lay["type"], "bc": lay["bc"]}, "puts": puts}
