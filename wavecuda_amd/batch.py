"""Batched universal-threshold smoothing (wh_batch_* of the C ABI).

Many signals of power-of-two lengths, mixed filters and mixed transforms
are smoothed (forward -> MAD universal threshold -> hard/soft threshold ->
inverse) by one persistent HIP launch fed by a device work queue.  Per
signal the result is what the single-signal smoothing computes with
threshtype UNIV over the full level range (reference src/rfunctions.cu:82-87
and src/thresh.cpp:170-201).

    plan = BatchPlan(lens, filters, ttypes, levels, dtype=torch.float64)
    y, t = plan.smooth(x)          # x: packed signals on the GPU
    plan.tune(x)                   # optional: time the launch shapes
    plan.capture(x, y, t); plan.replay()   # HIP graph of the pipeline
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as W


def _dtype_code(dtype) -> int:
    if dtype == torch.float64:
        return W.WH_F64
    if dtype == torch.float32:
        return W.WH_F32
    raise W.WaveHipError(f"unsupported dtype {dtype}")


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class BatchPlan:
    """Plan of a batch: signal table, workspace and launch shape on the
    current device."""

    def __init__(self, lens, filters, ttypes, levels, dtype=torch.float64, hardness=W.SOFT):
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        filters = np.ascontiguousarray(filters, dtype=np.int32)
        ttypes = np.ascontiguousarray(ttypes, dtype=np.int32)
        levels = np.ascontiguousarray(levels, dtype=np.int32)
        count = len(lens)
        if not (len(filters) == len(ttypes) == len(levels) == count):
            raise W.WaveHipError("signal tables differ in length")
        self.dtype = dtype
        self.count = count
        self.lens = lens
        self.offsets = np.concatenate([[0], np.cumsum(lens.astype(np.int64))])
        self._h = C.c_void_p()
        ip = C.POINTER(C.c_int)
        W.check(
            W.lib().wh_batch_create(
                lens.ctypes.data_as(ip), filters.ctypes.data_as(ip), ttypes.ctypes.data_as(ip),
                levels.ctypes.data_as(ip), count, _dtype_code(dtype), int(hardness),
                C.byref(self._h)),
            "wh_batch_create")
        self._keep = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                W.lib().wh_batch_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p()

    @property
    def total(self) -> int:
        return int(self.offsets[-1])

    def info(self) -> dict:
        tot, nt, wg, lds, ws, tuned = (C.c_int64(), C.c_int(), C.c_int(), C.c_int64(),
                                       C.c_int64(), C.c_int())
        W.check(W.lib().wh_batch_info(self._h, C.byref(tot), C.byref(nt), C.byref(wg),
                                      C.byref(lds), C.byref(ws), C.byref(tuned)),
                "wh_batch_info")
        return {"total": tot.value, "threads": nt.value, "workgroups": wg.value,
                "lds_bytes": lds.value, "workspace_bytes": ws.value, "tuned": bool(tuned.value)}

    def _check_x(self, x):
        if not (x.is_cuda and x.dtype == self.dtype and x.is_contiguous() and x.dim() == 1
                and x.numel() == self.total):
            raise W.WaveHipError("x must be a contiguous 1-D GPU tensor of the plan's dtype "
                                 "holding the packed signals")

    def _outs(self, x, y, thresh):
        if y is None:
            y = torch.empty_like(x)
        if thresh is None:
            thresh = torch.empty(self.count, dtype=torch.float64, device=x.device)
        return y, thresh

    def set_shape(self, threads: int, workgroups: int) -> None:
        """Launch shape as the tuner would set it (diagnostics: results never
        depend on the shape)."""
        W.check(W.lib().wh_debug_batch_set_shape(self._h, int(threads), int(workgroups)),
                "wh_debug_batch_set_shape")

    def smooth(self, x, y=None, thresh=None):
        """Smoothed signals (packed like x) and the per-signal thresholds."""
        self._check_x(x)
        y, thresh = self._outs(x, y, thresh)
        W.check(W.lib().wh_batch_smooth(self._h, _ptr(x), _ptr(y), _ptr(thresh), _stream()),
                "wh_batch_smooth")
        return y, thresh

    def tune(self, x, y=None, thresh=None):
        self._check_x(x)
        y, thresh = self._outs(x, y, thresh)
        W.check(W.lib().wh_batch_tune(self._h, _ptr(x), _ptr(y), _ptr(thresh), _stream()),
                "wh_batch_tune")
        return y, thresh

    def capture(self, x, y, thresh=None):
        """Capture the pipeline on these buffers into the plan's HIP graph."""
        self._check_x(x)
        torch.cuda.synchronize()
        W.check(W.lib().wh_batch_capture(self._h, _ptr(x), _ptr(y), _ptr(thresh)),
                "wh_batch_capture")
        self._keep = (x, y, thresh)  # the graph holds their addresses

    def replay(self):
        W.check(W.lib().wh_batch_replay(self._h, _stream()), "wh_batch_replay")

    def split(self, y):
        """List of per-signal views of a packed tensor."""
        o = self.offsets
        return [y[int(o[i]):int(o[i + 1])] for i in range(self.count)]
