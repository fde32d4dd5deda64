"""ALP / G-ALP / ALP-RD: host encoding through the library's C++ encoder
(include/flsgpu_host.h) and device decoding through the HIP kernels
(include/flsgpu.h). Field names follow src/flsgpu/structs.cuh."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi as A
from .device import _ptr, _stream_ptr, to_device


class AlpHostColumnStruct(C.Structure):
    _fields_ = [
        ("n_values", C.c_uint64), ("n_vecs", C.c_uint64), ("packed", C.c_void_p),
        ("bit_widths", C.c_void_p), ("vector_offsets", C.c_void_p), ("bases", C.c_void_p),
        ("factor_indices", C.c_void_p), ("fraction_indices", C.c_void_p),
        ("exceptions_offsets", C.c_void_p), ("exceptions", C.c_void_p), ("positions", C.c_void_p),
        ("counts", C.c_void_p), ("n_packed_words", C.c_uint64), ("n_exceptions", C.c_uint64),
        ("compressed_bytes_alp", C.c_uint64), ("compressed_bytes_galp", C.c_uint64),
    ]


class AlprdHostColumnStruct(C.Structure):
    _fields_ = [
        ("n_values", C.c_uint64), ("n_vecs", C.c_uint64), ("right_bit_width", C.c_uint32),
        ("left_bit_width", C.c_uint32), ("dictionary_size", C.c_uint32),
        ("dictionary", C.c_uint16 * 8), ("right_packed", C.c_void_p), ("left_packed", C.c_void_p),
        ("exceptions_offsets", C.c_void_p), ("exceptions", C.c_void_p), ("positions", C.c_void_p),
        ("counts", C.c_void_p), ("n_exceptions", C.c_uint64), ("compressed_bytes", C.c_uint64),
    ]


_P = C.c_void_p


def _enc_lib():
    L = A.lib()
    if not getattr(L, "_enc_sigs", False):
        for t in ("f32", "f64"):
            getattr(L, f"flsgpu_alp_encode_{t}").argtypes = [_P, C.c_uint64, C.POINTER(AlpHostColumnStruct), C.c_int]
            getattr(L, f"flsgpu_alp_encode_with_params_{t}").argtypes = [
                _P, C.c_uint64, _P, _P, C.POINTER(AlpHostColumnStruct), C.c_int]
            getattr(L, f"flsgpu_alp_search_{t}").argtypes = [_P, C.c_uint64, _P, _P, _P, C.c_int]
            getattr(L, f"flsgpu_galp_convert_{t}").argtypes = [C.c_uint64] + [_P] * 7
            getattr(L, f"flsgpu_alprd_encode_{t}").argtypes = [_P, C.c_uint64, C.POINTER(AlprdHostColumnStruct), C.c_int]
        for w in ("u32", "u64"):
            getattr(L, f"flsgpu_ffor_pack_{w}").argtypes = [_P, C.c_uint64, _P, _P, _P, _P]
        L.flsgpu_encoder_last_error.restype = C.c_char_p
        L.flsgpu_encoder_last_error.argtypes = []
        L._enc_sigs = True
    return L


def _check_enc(rc):
    if rc < 0:
        raise A.FlsgpuError(rc, _enc_lib().flsgpu_encoder_last_error().decode())


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclass
class AlpHostColumn:
    """host::ALPColumn<T>; offsets_counts set for the G-ALP (extended) layout."""
    n_values: int
    packed: np.ndarray
    bit_widths: np.ndarray
    vector_offsets: np.ndarray
    bases: np.ndarray
    factor_indices: np.ndarray
    fraction_indices: np.ndarray
    exceptions_offsets: np.ndarray
    exceptions: np.ndarray
    positions: np.ndarray
    counts: np.ndarray
    offsets_counts: np.ndarray | None = None
    compressed_bytes_alp: int = 0
    compressed_bytes_galp: int = 0

    @property
    def n_vecs(self):
        return self.n_values // 1024

    @property
    def dtype(self):
        return self.exceptions.dtype

    def compression_ratio(self, extended=False):
        return self.n_values * self.dtype.itemsize / (
            self.compressed_bytes_galp if extended else self.compressed_bytes_alp)


def _alloc_alp(n, dtype):
    ut = np.uint32 if dtype == np.float32 else np.uint64
    nv = n // 1024
    arrs = dict(packed=np.zeros(max(n, 1), ut), bit_widths=np.zeros(nv, np.uint8),
                vector_offsets=np.zeros(nv, np.uint64), bases=np.zeros(nv, ut),
                factor_indices=np.zeros(nv, np.uint8), fraction_indices=np.zeros(nv, np.uint8),
                exceptions_offsets=np.zeros(nv, np.uint64), exceptions=np.zeros(max(n, 1), dtype),
                positions=np.zeros(max(n, 1), np.uint16), counts=np.zeros(nv, np.uint16))
    st = AlpHostColumnStruct(0, 0, *[_np_ptr(arrs[k]) for k in (
        "packed", "bit_widths", "vector_offsets", "bases", "factor_indices", "fraction_indices",
        "exceptions_offsets", "exceptions", "positions", "counts")], 0, 0, 0, 0)
    return arrs, st


def _finish_alp(arrs, st, n):
    return AlpHostColumn(
        n, arrs["packed"][: st.n_packed_words].copy(), arrs["bit_widths"], arrs["vector_offsets"],
        arrs["bases"], arrs["factor_indices"], arrs["fraction_indices"],
        arrs["exceptions_offsets"], arrs["exceptions"][: st.n_exceptions].copy(),
        arrs["positions"][: st.n_exceptions].copy(), arrs["counts"], None,
        int(st.compressed_bytes_alp), int(st.compressed_bytes_galp))


def alp_encode(values: np.ndarray, n_threads: int = 0) -> AlpHostColumn | None:
    """alp::encode: None when sampling selects ALP_RD (the reference throws)."""
    values = np.ascontiguousarray(values)
    t = "f32" if values.dtype == np.float32 else "f64"
    arrs, st = _alloc_alp(len(values), values.dtype)
    rc = getattr(_enc_lib(), f"flsgpu_alp_encode_{t}")(_np_ptr(values), len(values), C.byref(st), n_threads)
    if rc == 1:
        return None
    _check_enc(rc)
    return _finish_alp(arrs, st, len(values))


def alp_encode_with_params(values: np.ndarray, exps: np.ndarray, facs: np.ndarray,
                           n_threads: int = 0) -> AlpHostColumn:
    values = np.ascontiguousarray(values)
    t = "f32" if values.dtype == np.float32 else "f64"
    arrs, st = _alloc_alp(len(values), values.dtype)
    e = np.ascontiguousarray(exps, dtype=np.uint8)
    f = np.ascontiguousarray(facs, dtype=np.uint8)
    _check_enc(getattr(_enc_lib(), f"flsgpu_alp_encode_with_params_{t}")(
        _np_ptr(values), len(values), _np_ptr(e), _np_ptr(f), C.byref(st), n_threads))
    return _finish_alp(arrs, st, len(values))


def alp_search(samples: np.ndarray, n_threads: int = 0):
    """Exhaustive per-vector (exponent, factor) search on [n_vecs, 32] samples."""
    samples = np.ascontiguousarray(samples)
    nv = samples.shape[0]
    t = "f32" if samples.dtype == np.float32 else "f64"
    e = np.zeros(nv, np.uint8)
    f = np.zeros(nv, np.uint8)
    est = np.zeros(nv, np.uint64)
    _check_enc(getattr(_enc_lib(), f"flsgpu_alp_search_{t}")(
        _np_ptr(samples), nv, _np_ptr(e), _np_ptr(f), _np_ptr(est), n_threads))
    return e, f, est


def galp_convert(col: AlpHostColumn) -> AlpHostColumn:
    """ALPColumn::create_extended_column (structs.cuh:308-331)."""
    t = "f32" if col.dtype == np.float32 else "f64"
    nl = 32 if col.dtype == np.float32 else 16
    exc = np.zeros(max(len(col.exceptions), 1), col.dtype)
    pos = np.zeros(max(len(col.positions), 1), np.uint16)
    oc = np.zeros(col.n_vecs * nl, np.uint16)
    src_e = col.exceptions if len(col.exceptions) else np.zeros(1, col.dtype)
    src_p = col.positions if len(col.positions) else np.zeros(1, np.uint16)
    _check_enc(getattr(_enc_lib(), f"flsgpu_galp_convert_{t}")(
        col.n_vecs, _np_ptr(src_e), _np_ptr(src_p), _np_ptr(col.counts),
        _np_ptr(col.exceptions_offsets), _np_ptr(exc), _np_ptr(pos), _np_ptr(oc)))
    n = len(col.exceptions)
    return AlpHostColumn(col.n_values, col.packed, col.bit_widths, col.vector_offsets, col.bases,
                         col.factor_indices, col.fraction_indices, col.exceptions_offsets,
                         exc[:n].copy(), pos[:n].copy(), col.counts, oc,
                         col.compressed_bytes_alp, col.compressed_bytes_galp)


# ---------------------------------------------------------------------------
# Device columns and decoding
# ---------------------------------------------------------------------------
@dataclass
class DeviceAlpColumn:
    n_values: int
    packed: torch.Tensor
    bit_widths: torch.Tensor
    vector_offsets: torch.Tensor
    bases: torch.Tensor
    factor_indices: torch.Tensor
    fraction_indices: torch.Tensor
    exceptions_offsets: torch.Tensor
    exceptions: torch.Tensor
    positions: torch.Tensor
    counts: torch.Tensor | None
    offsets_counts: torch.Tensor | None
    width: int

    @property
    def n_vecs(self):
        return self.n_values // 1024

    @property
    def extended(self):
        return self.offsets_counts is not None

    def struct(self) -> A.ALPColumn:
        bp = A.BPColumn(self.n_values, self.n_vecs, _ptr(self.packed), _ptr(self.bit_widths),
                        _ptr(self.vector_offsets))
        return A.ALPColumn(A.FFORColumn(bp, _ptr(self.bases)), _ptr(self.factor_indices),
                           _ptr(self.fraction_indices), int(self.exceptions.numel()),
                           _ptr(self.exceptions_offsets), _ptr(self.exceptions),
                           _ptr(self.positions), _ptr(self.counts), _ptr(self.offsets_counts))

    @staticmethod
    def from_host(col: AlpHostColumn, device="cuda") -> "DeviceAlpColumn":
        w = col.dtype.itemsize * 8
        return DeviceAlpColumn(
            col.n_values, to_device(col.packed, device), to_device(col.bit_widths, device),
            to_device(col.vector_offsets, device), to_device(col.bases, device),
            to_device(col.factor_indices, device), to_device(col.fraction_indices, device),
            to_device(col.exceptions_offsets, device), to_device(col.exceptions, device),
            to_device(col.positions, device),
            None if col.offsets_counts is not None else to_device(col.counts, device),
            None if col.offsets_counts is None else to_device(col.offsets_counts, device), w)


def _default_patcher(col: DeviceAlpColumn) -> int:
    return A.PATCHERS["prefetch-all"] if col.extended else A.PATCHERS["stateful"]


def decompress(col: DeviceAlpColumn, out: torch.Tensor | None = None, patcher: int | None = None,
               variant: int = 0, stream=None) -> torch.Tensor:
    if out is None:
        out = torch.empty(col.n_values, dtype=torch.float32 if col.width == 32 else torch.float64,
                          device=col.packed.device)
    patcher = _default_patcher(col) if patcher is None else patcher
    L = A.lib()
    sfx = "f32" if col.width == 32 else "f64"
    fn = getattr(L, f"flsgpu_decompress_alp_ext_{sfx}" if col.extended else f"flsgpu_decompress_alp_{sfx}")
    A.check(fn(C.byref(col.struct()), _ptr(out), patcher, variant, _stream_ptr(stream)))
    return out


def query_any(col: DeviceAlpColumn, magic: float, result: torch.Tensor | None = None,
              patcher: int | None = None, stream=None) -> torch.Tensor:
    if result is None:
        result = torch.empty(1, dtype=torch.uint8, device=col.packed.device)
    patcher = _default_patcher(col) if patcher is None else patcher
    sfx = "f32" if col.width == 32 else "f64"
    fn = getattr(A.lib(), f"flsgpu_query_any_alp_ext_{sfx}" if col.extended else f"flsgpu_query_any_alp_{sfx}")
    A.check(fn(C.byref(col.struct()), magic, _ptr(result), patcher, 0, _stream_ptr(stream)))
    return result


def query_multi_column(cols: list[DeviceAlpColumn], value: float, result: torch.Tensor | None = None,
                       patcher: int | None = None, stream=None) -> torch.Tensor:
    if result is None:
        result = torch.empty(1, dtype=torch.uint8, device=cols[0].packed.device)
    patcher = _default_patcher(cols[0]) if patcher is None else patcher
    arr = (A.ALPColumn * len(cols))(*[c.struct() for c in cols])
    sfx = "f32" if cols[0].width == 32 else "f64"
    fn = getattr(A.lib(), f"flsgpu_query_multi_column_alp_{sfx}")
    A.check(fn(arr, len(cols), value, _ptr(result), patcher, 0, _stream_ptr(stream)))
    return result


@dataclass
class AlprdHostColumn:
    """ALP-RD column in the flsgpu_alprd_column format (include/flsgpu.h)."""
    n_values: int
    right_bw: int
    left_bw: int
    dictionary: np.ndarray
    right_packed: np.ndarray
    left_packed: np.ndarray
    exceptions_offsets: np.ndarray
    exceptions: np.ndarray
    positions: np.ndarray
    counts: np.ndarray
    compressed_bytes: int = 0

    @property
    def n_vecs(self):
        return self.n_values // 1024

    @property
    def width(self):
        return self.right_packed.dtype.itemsize * 8


def alprd_encode(values: np.ndarray, n_threads: int = 0) -> AlprdHostColumn:
    values = np.ascontiguousarray(values)
    w = values.dtype.itemsize * 8
    t = "f32" if w == 32 else "f64"
    ut = np.uint32 if w == 32 else np.uint64
    n = len(values)
    nv = n // 1024
    rp = np.zeros(max(nv * (w - 1) * (1024 // w), 1), ut)
    lp = np.zeros(max(nv * 3 * 64, 1), np.uint16)
    eo = np.zeros(nv, np.uint64)
    exc = np.zeros(max(n, 1), np.uint16)
    pos = np.zeros(max(n, 1), np.uint16)
    cnt = np.zeros(nv, np.uint16)
    st = AlprdHostColumnStruct(0, 0, 0, 0, 0, (C.c_uint16 * 8)(), _np_ptr(rp), _np_ptr(lp),
                               _np_ptr(eo), _np_ptr(exc), _np_ptr(pos), _np_ptr(cnt), 0, 0)
    _check_enc(getattr(_enc_lib(), f"flsgpu_alprd_encode_{t}")(_np_ptr(values), n, C.byref(st), n_threads))
    rbw, lbw, ds = st.right_bit_width, st.left_bit_width, st.dictionary_size
    ne = int(st.n_exceptions)
    return AlprdHostColumn(n, rbw, lbw, np.array(list(st.dictionary)[:ds], np.uint16),
                           rp[: nv * rbw * (1024 // w)].copy(), lp[: nv * lbw * 64].copy(), eo,
                           exc[:ne].copy(), pos[:ne].copy(), cnt, int(st.compressed_bytes))


def ffor_pack(values: np.ndarray, bit_widths: np.ndarray, bases: np.ndarray | None):
    """FastLanes FFOR packing through the library (per-vector bit widths)."""
    values = np.ascontiguousarray(values)
    w = values.dtype.itemsize * 8
    nv = len(values) // 1024
    bws = np.ascontiguousarray(bit_widths, np.uint8)
    packed = np.zeros(max(int(bws.astype(np.int64).sum()) * (1024 // w), 1), values.dtype)
    offs = np.zeros(nv, np.uint64)
    b = None if bases is None else np.ascontiguousarray(bases, values.dtype)
    _check_enc(getattr(_enc_lib(), f"flsgpu_ffor_pack_u{w}")(
        _np_ptr(values), nv, _np_ptr(bws), _np_ptr(b), _np_ptr(packed), _np_ptr(offs)))
    return packed[: int(bws.astype(np.int64).sum()) * (1024 // w)], offs


@dataclass
class DeviceAlprdColumn:
    n_values: int
    right_bw: int
    left_bw: int
    dictionary: np.ndarray
    right_packed: torch.Tensor
    left_packed: torch.Tensor
    exceptions_offsets: torch.Tensor
    exceptions: torch.Tensor
    positions: torch.Tensor
    counts: torch.Tensor
    width: int

    @property
    def n_vecs(self):
        return self.n_values // 1024

    def struct(self) -> A.ALPRDColumn:
        d = (C.c_uint16 * 8)(*([int(x) for x in self.dictionary] + [0] * (8 - len(self.dictionary))))
        return A.ALPRDColumn(self.n_values, self.n_vecs, self.right_bw, self.left_bw,
                             len(self.dictionary), d, _ptr(self.right_packed), _ptr(self.left_packed),
                             _ptr(self.exceptions_offsets), _ptr(self.exceptions),
                             _ptr(self.positions), _ptr(self.counts))

    @staticmethod
    def from_host(col: AlprdHostColumn, device="cuda") -> "DeviceAlprdColumn":
        return DeviceAlprdColumn(col.n_values, col.right_bw, col.left_bw, col.dictionary,
                                 to_device(col.right_packed, device), to_device(col.left_packed, device),
                                 to_device(col.exceptions_offsets, device), to_device(col.exceptions, device),
                                 to_device(col.positions, device), to_device(col.counts, device), col.width)


def alprd_decompress(col: DeviceAlprdColumn, out: torch.Tensor | None = None, variant: int = 0,
                     stream=None) -> torch.Tensor:
    if out is None:
        out = torch.empty(col.n_values, dtype=torch.float32 if col.width == 32 else torch.float64,
                          device=col.right_packed.device)
    fn = getattr(A.lib(), f"flsgpu_decompress_alprd_{'f32' if col.width == 32 else 'f64'}")
    A.check(fn(C.byref(col.struct()), _ptr(out), variant, _stream_ptr(stream)))
    return out
