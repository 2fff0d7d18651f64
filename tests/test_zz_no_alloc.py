"""wh_modwt_inv_noalloc, the MODWT inverse that allocates nothing, against
wh_modwt_inv_many (the same kernels with library-allocated scratch): equal
bit for bit on two calls in a row, the second with a smaller input, both
with multi-group plans (so the static scratch carries the level hand-off);
sizes beyond its 16 MiB scratch are refused; the call leaves the device's
free memory unchanged."""
import ctypes as C

import pytest
import torch

from wavecuda_amd import _lib as W
from wavecuda_amd import device as D

pytestmark = pytest.mark.gpu


def _lib():
    L = W.lib()
    # set-up: wh_init places the static scratch on the device (the runtime
    # loads its code object), so no call below changes the free memory
    W.check(L.wh_init(-1), "wh_init")
    L.wh_modwt_inv_noalloc.restype = C.c_int
    L.wh_modwt_inv_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int,
                                       C.c_int, C.c_int, C.POINTER(W.wh_thresh), C.c_void_p]
    return L


CASES = [  # (n, count, J, filter, dtype, thresholded)
    (1 << 18, 4, 10, W.LA8, torch.float64, True),
    (1 << 14, 2, 8, W.D4, torch.float32, False),
]


def _run_pair(n, count, J, f, dt, thresholded):
    L = _lib()
    esz = W.WH_F64 if dt == torch.float64 else W.WH_F32
    plan = D.modwt_plan(True, n, J, f, dt, count=count)
    print(f"n={n} count={count} J={J} filter={f} {dt} plan {plan}", flush=True)
    assert plan.count("{") >= 2, "the case must hand levels between groups"
    g = torch.Generator(device="cuda").manual_seed(n + count)
    xmod = torch.randn(count, 2 * n * J, dtype=dt, device="cuda", generator=g)
    lam = torch.linspace(0.1, 0.6, J, dtype=torch.float64, device="cuda")
    thr = W.make_thresh(True, W.SOFT, 0, J - 1, 0.0, lam.data_ptr()) if thresholded else None
    tp = C.byref(thr) if thr is not None else None
    want = torch.empty(count, n, dtype=dt, device="cuda")
    W.check(L.wh_modwt_inv_many(xmod.data_ptr(), want.data_ptr(), n, count, J, f, esz, tp,
                                D._stream()), "wh_modwt_inv_many")
    got = torch.full((count, n), float("nan"), dtype=dt, device="cuda")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    W.check(L.wh_modwt_inv_noalloc(xmod.data_ptr(), got.data_ptr(), n, count, J, f, esz, tp,
                                   D._stream()), "wh_modwt_inv_noalloc")
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    assert torch.equal(got, want)


def test_noalloc_inverse_two_calls():
    for case in CASES:
        _run_pair(*case)


def test_noalloc_inverse_refuses_oversize():
    L = _lib()
    n, count, J = 1 << 20, 2, 10  # 2^21 FP64 samples: twice the static scratch
    a = torch.empty(2, dtype=torch.float64, device="cuda")
    rc = L.wh_modwt_inv_noalloc(a.data_ptr(), a.data_ptr(), n, count, J, W.LA8, W.WH_F64, None,
                                D._stream())
    assert rc == W.WH_ERR_ARG
    assert b"16 MiB" in L.wh_last_error()
