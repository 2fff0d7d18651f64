"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, where it was
built, of the reference's own CPU sources (oracle/_ref/libwvref.so).

Test infrastructure: used by tests/, the smoke check and bench.py's CPU
baseline only, as the checker.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libwvref.so"

_DP = C.POINTER(C.c_double)
_U = C.c_uint
_I = C.c_int

_o = None
_r = None


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def oracle():
    global _o
    if _o is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
        L = C.CDLL(str(ORACLE_SO))
        L.wvo_check_len_levels.restype = _U
        L.wvo_check_len_levels.argtypes = [_U, _U, _U]
        L.wvo_check_len_levels_mm.restype = _U
        L.wvo_check_len_levels_mm.argtypes = [_U, _U, _U, _U, _U]
        L.wvo_filter_coeffs.argtypes = [_I, _DP]
        L.wvo_dwt.argtypes = [_DP, _U, _I, _U, _I]
        L.wvo_dwt_omp.argtypes = [_DP, _U, _I, _U, _I]
        L.wvo_time_dwt_fwd_inv.restype = C.c_double
        L.wvo_time_dwt_fwd_inv.argtypes = [_DP, _DP, _U, _U, _I, _I]
        L.wvo_modwt.argtypes = [_DP, _DP, _U, _I, _U, _I]
        L.wvo_modwt_omp.argtypes = [_DP, _DP, _U, _I, _U, _I]
        L.wvo_thresh_coef.restype = C.c_double
        L.wvo_thresh_coef.argtypes = [C.c_double, C.c_double, _I]
        L.wvo_threshold.argtypes = [_DP, _DP, _U, _U, _I, _I, C.c_double, _I, _U, _U]
        L.wvo_median.restype = C.c_double
        L.wvo_median.argtypes = [_DP, _U]
        L.wvo_mad.restype = C.c_double
        L.wvo_mad.argtypes = [_DP, _U]
        L.wvo_univ_thresh.restype = C.c_double
        L.wvo_univ_thresh.argtypes = [_DP, _DP, _U, _U, _I, _U, _U]
        L.wvo_interp_mse.restype = C.c_double
        L.wvo_interp_mse.argtypes = [_DP, _DP, _DP, _U]
        L.wvo_cvt.restype = C.c_double
        L.wvo_cvt.argtypes = [_DP, _U, _U, _I, _I, _I, C.c_double, _U, _U, _DP, _I,
                              C.POINTER(_I), C.POINTER(_I)]
        L.wvo_smooth.argtypes = [_DP, _U, _U, _I, _I, _I, _DP, _I, _U, _U, C.c_double]
        L.wvo_modwt_convert.argtypes = [_DP, _DP, _U, _U, _I]
        L.wvo_modwt_po.argtypes = [_DP, _DP, _U, _I, _U, _I]
        L.wvo_coeff_positions.restype = C.c_long
        L.wvo_coeff_positions.argtypes = [C.c_long, _U, _I, _U, _I, C.POINTER(C.c_long)]
        L.wvo_coeff_labels.argtypes = [C.c_long, _U, _I, C.POINTER(_I), _DP,
                                       C.POINTER(C.c_ubyte)]
        _o = L
    return _o


def ref_available() -> bool:
    return REF_SO.exists()


def ref():
    global _r
    if _r is None:
        L = C.CDLL(str(REF_SO))
        L.ref_transform.argtypes = [_DP, _DP, _I, _I, _I, _I, _I, _I]
        L.ref_threshold.argtypes = [_DP, _DP, _I, _I, _I, _I, _I, C.c_double, _I, _I, _I]
        L.ref_smooth.argtypes = [_DP, _I, _I, _I, _I, _I, _I, _DP, _I, _I, _I, C.c_double]
        L.ref_median.restype = C.c_double
        L.ref_median.argtypes = [_DP, _I]
        L.ref_mad.restype = C.c_double
        L.ref_mad.argtypes = [_DP, _I]
        L.ref_univ_thresh.restype = C.c_double
        L.ref_univ_thresh.argtypes = [_DP, _DP, _I, _I, _I, _I, _I, _I, _I]
        L.ref_interp_mse.restype = C.c_double
        L.ref_interp_mse.argtypes = [_DP, _DP, _DP, _I]
        L.ref_check_len_levels.restype = _U
        L.ref_check_len_levels.argtypes = [_U, _U, _U]
        L.ref_convert_modwt.argtypes = [_DP, _DP, _I, _I, _I]
        _r = L
    return _r


FLEN = {1: 2, 2: 4, 3: 6, 4: 8}


def levels(n, nlevels, filt):
    return oracle().wvo_check_len_levels(n, nlevels, FLEN[filt])


# ---- numpy conveniences over the oracle ---------------------------------

def dwt(x, nlevels, filt, sense=1, omp=False):
    y = np.array(x, dtype=np.float64, copy=True)
    fn = oracle().wvo_dwt_omp if omp else oracle().wvo_dwt
    rc = fn(_p(y), len(y), sense, nlevels, filt)
    assert rc == 0, rc
    return y


def modwt_fwd(x, nlevels, filt, omp=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    J = levels(n, nlevels, filt)
    xm = np.zeros(2 * n * J)
    fn = oracle().wvo_modwt_omp if omp else oracle().wvo_modwt
    rc = fn(_p(x), _p(xm), n, 1, J, filt)
    assert rc == 0, rc
    return xm


def modwt_inv(xmod, n, nlevels, filt, omp=False):
    J = levels(n, nlevels, filt)
    xm = np.array(xmod[: 2 * n * J], dtype=np.float64, copy=True)
    x = np.zeros(n)
    fn = oracle().wvo_modwt_omp if omp else oracle().wvo_modwt
    rc = fn(_p(x), _p(xm), n, 0, J, filt)
    assert rc == 0, rc
    return x


def threshold(coeffs, n, nlevels, ttype, filt, t, hardness, minlevel, maxlevel):
    c = np.array(coeffs, dtype=np.float64, copy=True)
    rc = oracle().wvo_threshold(_p(c), _p(c), n, nlevels, ttype, filt, t, hardness, minlevel,
                                maxlevel)
    assert rc == 0, rc
    return c


def threshold_levels_(c, n, nlevels, ttype, filt, lams, hardness):
    """In place: the details of 0-based level l thresholded at lams[l], level
    by level through the oracle's rule (src/thresh.cpp:15-83, :142-157)."""
    assert c.dtype == np.float64 and c.flags.c_contiguous
    for l, t in enumerate(lams):
        rc = oracle().wvo_threshold(_p(c), _p(c), n, nlevels, ttype, filt, float(t), hardness,
                                    l, l)
        assert rc == 0, rc
    return c


def median(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return oracle().wvo_median(_p(x), len(x))


def mad(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return oracle().wvo_mad(_p(x), len(x))


def univ_thresh(coeffs, n, nlevels, ttype, minlevel, maxlevel):
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    if ttype == 0:
        return oracle().wvo_univ_thresh(_p(c), None, n, nlevels, 0, minlevel, maxlevel)
    return oracle().wvo_univ_thresh(None, _p(c), n, nlevels, 1, minlevel, maxlevel)


def interp_mse(xn, ye, yo):
    xn, ye, yo = (np.ascontiguousarray(a, dtype=np.float64) for a in (xn, ye, yo))
    return oracle().wvo_interp_mse(_p(xn), _p(ye), _p(yo), len(xn))


def cvt(x, nlevels, ttype, filt, hardness, tol, minlevel, maxlevel, trace_cap=64):
    y = np.array(x, dtype=np.float64, copy=True)
    tr = np.zeros(6 * trace_cap)
    nt = C.c_int(0)
    st = C.c_int(0)
    t = oracle().wvo_cvt(_p(y), len(y), nlevels, ttype, filt, hardness, tol, minlevel, maxlevel,
                         _p(tr), trace_cap, C.byref(nt), C.byref(st))
    return t, y, tr[: 6 * nt.value].reshape(-1, 6), st.value


def smooth(x, nlevels, ttype, filt, threshtype, thresh, hardness, minlevel, maxlevel, tol=0.01):
    y = np.array(x, dtype=np.float64, copy=True)
    t = np.array([thresh if thresh is not None else 0.0])
    rc = oracle().wvo_smooth(_p(y), len(y), nlevels, ttype, filt, threshtype, _p(t), hardness,
                             minlevel, maxlevel, tol)
    return y, float(t[0]), rc


def modwt_convert(xmod, n, nlevels, intype):
    """convert_modwt: intype 1 = time ordered in, 2 = packet ordered in."""
    a = np.ascontiguousarray(xmod, dtype=np.float64)
    out = np.zeros_like(a)
    assert oracle().wvo_modwt_convert(_p(a), _p(out), n, nlevels, intype) == 0
    return out


def modwt_po_fwd(x, nlevels, filt):
    n = x.size
    J = levels(n, nlevels, filt)
    xm = np.zeros(2 * n * J)
    assert oracle().wvo_modwt_po(_p(np.ascontiguousarray(x, dtype=np.float64).copy()), _p(xm),
                                 n, 1, J, filt) == 0
    return xm


def modwt_po_inv(xmod, n, nlevels, filt):
    J = levels(n, nlevels, filt)
    x = np.zeros(n)
    assert oracle().wvo_modwt_po(_p(x), _p(np.ascontiguousarray(xmod, dtype=np.float64).copy()),
                                 n, 0, J, filt) == 0
    return x


def coeff_positions(n, nlevels, ttype, level, coeff_type="d"):
    """getCoeffLevel's 1-based positions (R/rplotting.R:10-46); None where
    the R code stops."""
    pos = np.zeros(max(n, 1), dtype=np.int64)
    k = oracle().wvo_coeff_positions(n, nlevels, ttype, level, 1 if coeff_type == "s" else 0,
                                     pos.ctypes.data_as(C.POINTER(C.c_long)))
    return pos[:k] if k > 0 else None


def coeff_labels(n, nlevels, ttype):
    """WSTtoDT's (level, translate, is_scaling) columns in storage order."""
    m = n if ttype == 0 else 2 * n * nlevels
    level = np.zeros(m, dtype=np.int32)
    translate = np.zeros(m, dtype=np.float64)
    kind = np.zeros(m, dtype=np.uint8)
    rc = oracle().wvo_coeff_labels(n, nlevels, ttype, level.ctypes.data_as(C.POINTER(_I)),
                                   _p(translate), kind.ctypes.data_as(C.POINTER(C.c_ubyte)))
    assert rc == 0
    return level, translate, kind


def relerr(a, b):
    """Norm-wise relative error max|a-b| / max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / den)
