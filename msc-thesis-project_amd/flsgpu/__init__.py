"""flsgpu: MI355X (gfx950) FastLanes / ALP / ALP-RD decoder.

Host-side Python face of the C ABI in include/flsgpu.h. All decoding runs in
the HIP kernels of lib/libflsgpu_hip.so; PyTorch only provides device memory
and streams. Nothing here falls back to a CPU decoder.
"""
from . import _capi
from ._capi import FlsgpuError, lib, VARIANTS, PATCHERS, PRED_RANGE, PRED_OFFSET_LOWER_HALF, PRED_ALL

__all__ = ["_capi", "FlsgpuError", "lib", "VARIANTS", "PATCHERS", "PRED_RANGE",
           "PRED_OFFSET_LOWER_HALF", "PRED_ALL"]
