"""wavecuda_amd: MI355X-native wavelet engine (DWT / MODWT / thresholding /
cross-validated smoothing) behind the C ABI of libwavehip.so.

Submodules:
  _lib     ctypes binding of the C ABI (include/wavehip.h)
  device   tensor-level API on GPU tensors
  rsurface the reference package's R functions (GPUTransform, GPUReconstruct,
           GPUSmooth, ...) mirrored in Python with the same WST fields
  signals  synthetic doppler / white-noise inputs
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
