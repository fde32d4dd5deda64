"""Device-resident columns and the decode / query calls, over the C ABI.

Device memory and streams come from PyTorch (plumbing only); every decode and
query runs in the HIP kernels of libflsgpu_hip.so. Field names follow the
reference's column structs (src/flsgpu/structs.cuh:33-84).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _capi as A

_TORCH_OF = {
    np.dtype(np.uint8): torch.uint8,
    np.dtype(np.uint16): torch.uint16,
    np.dtype(np.uint32): torch.uint32,
    np.dtype(np.uint64): torch.uint64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float64): torch.float64,
}


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def to_device(a: np.ndarray, device="cuda") -> torch.Tensor:
    """Upload a numpy array as a 1-D device tensor. An empty array becomes a
    one-element zero tensor, so every column field has a valid device address
    (e.g. the packed array of a column whose vectors all have bit width 0)."""
    a = np.ascontiguousarray(a)
    if a.size == 0:
        a = np.zeros(1, dtype=a.dtype)
    t = torch.from_numpy(a.view(np.uint8).copy()).to(device)
    return t.view(_TORCH_OF[a.dtype])


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


@dataclass
class DeviceFFORColumn:
    """device::FFORColumn<T> (bases=None for device::BPColumn<T>)."""
    n_values: int
    packed: torch.Tensor
    bit_widths: torch.Tensor
    vector_offsets: torch.Tensor
    bases: torch.Tensor | None
    width: int = 32

    @property
    def n_vecs(self) -> int:
        return self.n_values // 1024

    def bp_struct(self) -> A.BPColumn:
        return A.BPColumn(self.n_values, self.n_vecs, _ptr(self.packed), _ptr(self.bit_widths),
                          _ptr(self.vector_offsets))

    def ffor_struct(self) -> A.FFORColumn:
        return A.FFORColumn(self.bp_struct(), _ptr(self.bases))

    @staticmethod
    def from_host(packed: np.ndarray, bit_widths: np.ndarray, vector_offsets: np.ndarray,
                  bases: np.ndarray | None, n_values: int, device="cuda") -> "DeviceFFORColumn":
        w = packed.dtype.itemsize * 8
        return DeviceFFORColumn(
            n_values, to_device(packed, device), to_device(bit_widths.astype(np.uint8), device),
            to_device(vector_offsets.astype(np.uint64), device),
            None if bases is None else to_device(bases.astype(packed.dtype), device), w)


def _out_dtype(width: int, float_out=False):
    if float_out:
        return torch.float32 if width == 32 else torch.float64
    return torch.uint32 if width == 32 else torch.uint64


def decompress(col: DeviceFFORColumn, out: torch.Tensor | None = None, variant: int = 0,
               stream=None) -> torch.Tensor:
    """bindings::decompress_column for BP / FFOR columns."""
    if out is None:
        out = torch.empty(col.n_values, dtype=_out_dtype(col.width), device=col.packed.device)
    s = _stream_ptr(stream)
    L = A.lib()
    if col.bases is None:
        fn = L.flsgpu_decompress_bp_u32 if col.width == 32 else L.flsgpu_decompress_bp_u64
        A.check(fn(C.byref(col.bp_struct()), _ptr(out), variant, s))
    else:
        fn = L.flsgpu_decompress_ffor_u32 if col.width == 32 else L.flsgpu_decompress_ffor_u64
        A.check(fn(C.byref(col.ffor_struct()), _ptr(out), variant, s))
    return out


def query_any(col: DeviceFFORColumn, magic: int, result: torch.Tensor | None = None,
              variant: int = 0, stream=None) -> torch.Tensor:
    """bindings::query_column: 1-byte device tensor, 1 iff any value == magic."""
    if result is None:
        result = torch.empty(1, dtype=torch.uint8, device=col.packed.device)
    L = A.lib()
    s = _stream_ptr(stream)
    if col.bases is None:
        fn = L.flsgpu_query_any_bp_u32 if col.width == 32 else L.flsgpu_query_any_bp_u64
        A.check(fn(C.byref(col.bp_struct()), int(magic), _ptr(result), variant, s))
    else:
        fn = L.flsgpu_query_any_ffor_u32 if col.width == 32 else L.flsgpu_query_any_ffor_u64
        A.check(fn(C.byref(col.ffor_struct()), int(magic), _ptr(result), variant, s))
    return result


def query_filter_sum(col: DeviceFFORColumn, predicate: int = A.PRED_OFFSET_LOWER_HALF,
                     lo: int = 0, hi: int = 0, result: torch.Tensor | None = None,
                     variant: int = 0, stream=None) -> torch.Tensor:
    """Fused decode + filter + sum: device uint64 tensor [count, sum]."""
    if result is None:
        result = torch.empty(2, dtype=torch.uint64, device=col.packed.device)
    L = A.lib()
    kind = "bp" if col.bases is None else "ffor"
    fn = getattr(L, f"flsgpu_query_filter_sum_{kind}_u{col.width}")
    desc = col.bp_struct() if col.bases is None else col.ffor_struct()
    A.check(fn(C.byref(desc), predicate, int(lo), int(hi), _ptr(result), variant,
               _stream_ptr(stream)))
    return result


def compute(col: DeviceFFORColumn, n_repetitions: int = 10, result: torch.Tensor | None = None,
            variant: int = 0, stream=None) -> torch.Tensor:
    """bindings::compute_column (synthetic ALU load + any == 1)."""
    if result is None:
        result = torch.empty(1, dtype=torch.uint8, device=col.packed.device)
    L = A.lib()
    kind = "bp" if col.bases is None else "ffor"
    fn = getattr(L, f"flsgpu_compute_{kind}_u{col.width}")
    desc = col.bp_struct() if col.bases is None else col.ffor_struct()
    A.check(fn(C.byref(desc), n_repetitions, _ptr(result), variant, _stream_ptr(stream)))
    return result


def query_multi_column(cols: list[DeviceFFORColumn], value: int,
                       result: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """bindings::query_multi_column over 1..10 FFOR columns."""
    if result is None:
        result = torch.empty(1, dtype=torch.uint8, device=cols[0].packed.device)
    arr = (A.FFORColumn * len(cols))(*[c.ffor_struct() for c in cols])
    L = A.lib()
    fn = (L.flsgpu_query_multi_column_ffor_u32 if cols[0].width == 32
          else L.flsgpu_query_multi_column_ffor_u64)
    A.check(fn(arr, len(cols), int(value), _ptr(result), 0, _stream_ptr(stream)))
    return result
