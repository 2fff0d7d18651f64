"""ctypes binding of libwavehip.so (the C ABI declared in include/wavehip.h).

The library is loaded from this package directory only.  There is no
fallback: if it is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libwavehip.so"

WH_OK = 0
WH_ERR_ARG, WH_ERR_HIP, WH_ERR_NOMEM, WH_ERR_UNSUPPORTED, WH_ERR_NODEVICE = 1, 2, 3, 4, 5
WH_F64 = 0
WH_F32 = 1

FWD, BWD = 1, 0
DWT, MODWT, MODWT_PO = 0, 1, 2
HAAR, D4, C6, LA8 = 1, 2, 3, 4
HARD, SOFT = 0, 1
MANUAL, UNIV, CV = 0, 1, 2

FILTER_CODES = {"Haar": HAAR, "D4": D4, "C6": C6, "LA8": LA8}
FILTER_NAMES = {v: k for k, v in FILTER_CODES.items()}
FILTER_LENGTHS = {HAAR: 2, D4: 4, C6: 6, LA8: 8}


class WaveHipError(RuntimeError):
    """Raised when a wh_* call returns a non-zero status."""


class wh_thresh(C.Structure):
    _fields_ = [
        ("enabled", C.c_int),
        ("hardness", C.c_int),
        ("minlevel", C.c_int),
        ("maxlevel", C.c_int),
        ("value", C.c_double),
        ("levels", C.c_void_p),
    ]


_VP = C.c_void_p
_I64 = C.c_int64
_I = C.c_int
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "wh_last_error": (C.c_char_p, []),
    "wh_version": (_I, []),
    "wh_init": (_I, [_I]),
    "wh_resolve_levels": (_I, [_I64, _I, _I]),
    "wh_modwt_fwd": (_I, [_VP, _VP, _I64, _I, _I, _I, _VP]),
    "wh_modwt_inv": (_I, [_VP, _VP, _I64, _I, _I, _I, C.POINTER(wh_thresh), _VP]),
    "wh_dwt_fwd": (_I, [_VP, _VP, _I64, _I, _I, _I, _VP]),
    "wh_dwt_inv": (_I, [_VP, _VP, _I64, _I, _I, _I, C.POINTER(wh_thresh), _VP]),
    "wh_threshold": (_I, [_VP, _I64, _I, _I, _I, _I, C.POINTER(wh_thresh), _VP]),
    "wh_median": (_I, [_VP, _I64, _I, _VP, _VP]),
    "wh_mad": (_I, [_VP, _I64, _I, _VP, _VP]),
    "wh_univ_thresh": (_I, [_VP, _I64, _I, _I, _I, _I, _I, _I, _VP, _VP]),
    "wh_interp_mse": (_I, [_VP, _VP, _VP, _I64, _I, _VP, _VP]),
    "wh_smooth": (_I, [_VP, _I64, _I, _I, _I, _I, _I, _DP, _I, _I, _I, C.c_double, _VP]),
    "RgpuTransform": (None, [_DP, _DP, _IP, _IP, _IP, _IP, _IP, _IP]),
    "RcpuTransform": (None, [_DP, _DP, _IP, _IP, _IP, _IP, _IP, _IP]),
    "RcpuThreshold": (None, [_DP, _DP, _IP, _IP, _IP, _IP, _IP, _DP, _IP, _IP, _IP]),
    "RgpuSmooth": (None, [_DP, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _IP, _IP, _IP, _DP]),
    "RcpuSmooth": (None, [_DP, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _IP, _IP, _IP, _DP]),
    "RgpuSmoothList": (None, [_DP, _IP, _IP, _IP, _IP, _IP, _IP, _DP]),
    "wh_debug_modwt_plan": (_I, [_I, _I64, _I, _I, _I, C.c_char_p, _I]),
    "wh_debug_modwt_plan_many": (_I, [_I, _I64, _I64, _I, _I, _I, C.c_char_p, _I]),
    "wh_debug_modwt_set_plan": (_I, [_I, _I64, _I, _I, _I, _I, _I, _I, _I]),
    "wh_modwt_tune": (_I, [_I64, _I, _I, _I, _VP]),
    "wh_coeff_level": (_I, [_VP, _VP, _I64, _I, _I, _I, _I, _I, _I, C.POINTER(_I64), _VP]),
    "wh_modwt_convert": (_I, [_VP, _VP, _I64, _I, _I, _I, _I, _VP]),
    "wh_coeff_labels": (_I, [_I64, _I, _I, _I, _VP, _VP, _VP, _VP]),
    "wh_modwt_fwd_many": (_I, [_VP, _VP, _I64, _I64, _I, _I, _I, _VP]),
    "wh_modwt_inv_many": (_I, [_VP, _VP, _I64, _I64, _I, _I, _I, C.POINTER(wh_thresh), _VP]),
    "wh_modwt_fwd_adj": (_I, [_VP, _VP, _I64, _I64, _I, _I, _I, _VP]),
    "wh_modwt_inv_adj": (_I, [_VP, _VP, _I64, _I64, _I, _I, _I, _VP]),
    "wh_dwt_fwd_adj": (_I, [_VP, _VP, _I64, _I, _I, _I, _VP]),
    "wh_dwt_inv_adj": (_I, [_VP, _VP, _I64, _I, _I, _I, _VP]),
    "wh_threshold_many": (_I, [_VP, _I64, _I64, _I, _I, _I, _I, C.POINTER(wh_thresh), _VP]),
    "wh_thresh_grad": (_I, [_VP, _VP, _VP, _VP, _I64, _I64, _I, _I, _I, _I,
                            C.POINTER(wh_thresh), _VP]),
    "wh_batch_create": (_I, [_IP, _IP, _IP, _IP, _I, _I, _I, C.POINTER(_VP)]),
    "wh_batch_destroy": (None, [_VP]),
    "wh_batch_smooth": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "wh_batch_tune": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "wh_batch_capture": (_I, [_VP, _VP, _VP, _VP]),
    "wh_batch_replay": (_I, [_VP, _VP]),
    "wh_debug_batch_set_shape": (_I, [_VP, _I, _I]),
    "wh_batch_info": (_I, [_VP, C.POINTER(_I64), _IP, _IP, C.POINTER(_I64), C.POINTER(_I64),
                           _IP]),
}

class wh_wt_entry(C.Structure):
    _fields_ = [
        ("wt_level", C.c_int),
        ("coeff_type", C.c_int),
        ("offset", C.c_int64),
        ("count", C.c_int64),
    ]


SIGNATURES["wh_wavethresh_layout"] = (_I, [_I64, _I, _I, _I, C.POINTER(wh_wt_entry), _I, _IP,
                                            C.c_char_p, _I, _IP])
SIGNATURES["wh_wavethresh_pack"] = (_I, [_VP, _VP, _I64, _I, _I, _I, _I, _VP])

_lib = None


def lib() -> C.CDLL:
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise WaveHipError(
                f"{LIB_PATH} is missing: build it with `python -m wavecuda_amd.build`"
            )
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def has(name: str) -> bool:
    return getattr(lib(), name, None) is not None


def check(rc: int, what: str = "") -> None:
    if rc != WH_OK:
        msg = lib().wh_last_error().decode(errors="replace")
        raise WaveHipError(f"{what or 'wavehip'} failed (status {rc}): {msg}")


def make_thresh(enabled=False, hardness=SOFT, minlevel=0, maxlevel=0, value=0.0, levels_ptr=None):
    t = wh_thresh()
    t.enabled = 1 if enabled else 0
    t.hardness = int(hardness)
    t.minlevel = int(minlevel)
    t.maxlevel = int(maxlevel)
    t.value = float(value)
    t.levels = levels_ptr
    return t
