"""The C-ABI library builds, loads without a GPU and exports every function
declared in include/wavehip.h (no compute calls here)."""
import re
from pathlib import Path

import pytest

from wavecuda_amd import _lib as W

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "wavehip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted({n for n in names if n not in {"if", "while", "for", "return", "sizeof"}})


def test_header_lists_entry_points():
    names = declared_functions()
    for must in ("RgpuTransform", "RgpuSmooth", "wh_modwt_fwd", "wh_modwt_inv", "wh_dwt_fwd",
                 "wh_dwt_inv", "wh_smooth", "wh_mad", "wh_univ_thresh"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    assert W.has(name), f"{name} declared in include/wavehip.h but not exported"


def test_no_device_needed_for_pure_queries():
    L = W.lib()
    assert L.wh_version() >= 100
    assert L.wh_resolve_levels(1024, 0, 2) == 9
    assert L.wh_resolve_levels(1024, 0, 4) == 8
    assert L.wh_resolve_levels(1000, 0, 1) == 0
    assert L.wh_resolve_levels(256, 7, 4) == 0
    assert L.wh_resolve_levels(8, 0, 3) == 1
    assert isinstance(L.wh_last_error(), bytes)


def test_argument_checks_before_any_device_call():
    """Argument errors come back as WH_ERR_ARG before the library touches a
    device: null pointers, codes out of range, and buffers not aligned to
    two elements (the kernels move element pairs with one vector access)."""
    import ctypes as C

    L = W.lib()
    good, odd8, odd4 = C.c_void_p(1 << 20), C.c_void_p((1 << 20) + 8), C.c_void_p((1 << 20) + 4)
    assert L.wh_modwt_fwd(None, good, 1024, 4, 4, 0, None) == 1
    assert L.wh_modwt_fwd(good, good, 1024, 4, 9, 0, None) == 1
    assert L.wh_modwt_fwd(good, good, 1000, 4, 4, 0, None) == 1
    assert L.wh_modwt_fwd(good, good, 1024, 4, 4, 7, None) == 1
    assert L.wh_modwt_fwd(odd8, good, 1024, 4, 4, 0, None) == 1  # FP64 pairs: 16 bytes
    assert b"aligned" in L.wh_last_error()
    assert L.wh_dwt_inv(good, odd4, 1024, 4, 2, 1, None, None) == 1  # FP32 pairs: 8 bytes
    assert b"aligned" in L.wh_last_error()
