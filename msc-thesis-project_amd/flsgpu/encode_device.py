"""Device-side encoding (include/flsgpu.h, flsgpu_*_encode_device_*): ALP
columns from per-vector (exponent, factor) and FFOR columns, built in HBM
from values already on the GPU. The bytes equal the host encoder's."""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi as A
from .alp import DeviceAlpColumn
from .device import DeviceFFORColumn, _stream_ptr

_P = C.c_void_p
_UT = {32: torch.uint32, 64: torch.uint64}


def _sigs():
    L = A.lib()
    if not getattr(L, "_device_encode_sigs", False):
        L.flsgpu_encode_workspace_bytes.argtypes = [C.c_uint64]
        L.flsgpu_encode_workspace_bytes.restype = C.c_size_t
        for t in ("f32", "f64"):
            f = getattr(L, f"flsgpu_alp_encode_device_{t}")
            f.argtypes = [_P, C.c_uint64, _P, _P, C.c_int, C.POINTER(A.ALPColumn), _P, _P, _P]
            f.restype = C.c_int
        for t in ("u32", "u64"):
            f = getattr(L, f"flsgpu_ffor_encode_device_{t}")
            f.argtypes = [_P, C.c_uint64, C.POINTER(A.FFORColumn), _P, _P, _P]
            f.restype = C.c_int
        L._device_encode_sigs = True
    return L


def _workspace(n_vecs: int, device) -> torch.Tensor:
    nbytes = int(_sigs().flsgpu_encode_workspace_bytes(n_vecs))
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def alp_encode_device(values: torch.Tensor, exps: torch.Tensor, facs: torch.Tensor,
                      galp: bool = False, stream=None):
    """Encode a device column of float32/float64 (n % 1024 == 0) with the
    given per-vector exponent/factor (uint8 device tensors). Returns
    (DeviceAlpColumn, {"packed_words", "exceptions"}); synchronises once to
    read the sizes."""
    assert values.dtype in (torch.float32, torch.float64) and values.numel() % 1024 == 0
    dev = values.device
    w = 32 if values.dtype == torch.float32 else 64
    n = values.numel()
    nv = n // 1024
    nl = 1024 // w
    packed = torch.empty(n, dtype=_UT[w], device=dev)
    bit_widths = torch.empty(nv, dtype=torch.uint8, device=dev)
    vector_offsets = torch.empty(nv, dtype=torch.uint64, device=dev)
    bases = torch.empty(nv, dtype=_UT[w], device=dev)
    fac_idx = torch.empty(nv, dtype=torch.uint8, device=dev)
    frac_idx = torch.empty(nv, dtype=torch.uint8, device=dev)
    exc_off = torch.empty(nv, dtype=torch.uint64, device=dev)
    exceptions = torch.empty(n, dtype=values.dtype, device=dev)
    positions = torch.empty(n, dtype=torch.uint16, device=dev)
    counts = torch.empty(nv, dtype=torch.uint16, device=dev)
    offsets_counts = torch.empty(nv * nl, dtype=torch.uint16, device=dev) if galp else None
    totals = torch.zeros(3, dtype=torch.int64, device=dev)
    ws = _workspace(nv, dev)
    ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
    bp = A.BPColumn(n, nv, ptr(packed), ptr(bit_widths), ptr(vector_offsets))
    col = A.ALPColumn(A.FFORColumn(bp, ptr(bases)), ptr(fac_idx), ptr(frac_idx), 0, ptr(exc_off),
                      ptr(exceptions), ptr(positions), ptr(counts), ptr(offsets_counts))
    fn = getattr(_sigs(), f"flsgpu_alp_encode_device_f{w}")
    A.check(fn(values.data_ptr(), nv, exps.data_ptr(), facs.data_ptr(), int(galp), C.byref(col),
               totals.data_ptr(), ws.data_ptr(), _stream_ptr(stream)))
    words, n_exc, overflow = (int(x) for x in totals.cpu().tolist())
    if galp and overflow:
        raise A.FlsgpuError(A.FLSGPU_EINVAL, "a lane holds more than 63 exceptions; "
                            "the G-ALP header cannot encode it")
    d = DeviceAlpColumn(n, packed[:max(words, 1)], bit_widths, vector_offsets, bases, fac_idx,
                        frac_idx, exc_off, exceptions[:max(n_exc, 1)], positions[:max(n_exc, 1)],
                        None if galp else counts, offsets_counts, w)
    return d, {"packed_words": words, "exceptions": n_exc, "counts": counts}


def ffor_encode_device(values: torch.Tensor, stream=None):
    """FFOR-encode a device column of uint32/uint64 (n % 1024 == 0): base =
    vector minimum, width of (max - min). Returns (DeviceFFORColumn, words)."""
    assert values.dtype in (torch.uint32, torch.uint64) and values.numel() % 1024 == 0
    dev = values.device
    w = 32 if values.dtype == torch.uint32 else 64
    n = values.numel()
    nv = n // 1024
    packed = torch.empty(n, dtype=values.dtype, device=dev)
    bit_widths = torch.empty(nv, dtype=torch.uint8, device=dev)
    vector_offsets = torch.empty(nv, dtype=torch.uint64, device=dev)
    bases = torch.empty(nv, dtype=values.dtype, device=dev)
    totals = torch.zeros(3, dtype=torch.int64, device=dev)
    ws = _workspace(nv, dev)
    bp = A.BPColumn(n, nv, packed.data_ptr(), bit_widths.data_ptr(), vector_offsets.data_ptr())
    col = A.FFORColumn(bp, bases.data_ptr())
    fn = getattr(_sigs(), f"flsgpu_ffor_encode_device_u{w}")
    A.check(fn(values.data_ptr(), nv, C.byref(col), totals.data_ptr(), ws.data_ptr(),
               _stream_ptr(stream)))
    words = int(totals[0].item())
    return DeviceFFORColumn(n, packed[:max(words, 1)], bit_widths, vector_offsets, bases, w), words
