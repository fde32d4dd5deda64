"""Column files (.flsc) through the library (include/flsgpu.h,
flsgpu_file_*): write a host column, read it back on the host, or load it
into one device block and decode / query it there; and uploads of host
columns into one device block (flsgpu_column_upload / flsgpu_column_free).

Host columns are the encoder's outputs: AlpHostColumn (standard layout, or
G-ALP when offsets_counts is set), AlprdHostColumn, and FforHostColumn below
for bit-packed / FFOR integers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi as A
from .alp import AlpHostColumn, AlprdHostColumn

_P = C.c_void_p
_UT = {32: np.uint32, 64: np.uint64}
_FT = {32: np.float32, 64: np.float64}

ENC = {"bp_u32": 0, "ffor_u32": 1, "bp_u64": 2, "ffor_u64": 3, "alp_f32": 4, "alp_f64": 5,
       "alp_ext_f32": 6, "alp_ext_f64": 7, "alprd_f32": 8, "alprd_f64": 9}
ENC_NAME = {v: k for k, v in ENC.items()}


class FileInfo(C.Structure):
    _fields_ = [("encoding", C.c_int), ("n_values", C.c_uint64), ("n_vecs", C.c_uint64),
                ("payload_bytes", C.c_uint64), ("file_bytes", C.c_uint64)]


def _lib():
    L = A.lib()
    if not getattr(L, "_file_sigs", False):
        L.flsgpu_file_write.argtypes = [C.c_char_p, C.c_int, _P]
        L.flsgpu_file_write.restype = C.c_int
        L.flsgpu_file_stat.argtypes = [C.c_char_p, C.POINTER(FileInfo)]
        L.flsgpu_file_stat.restype = C.c_int
        L.flsgpu_file_load.argtypes = [C.c_char_p, C.c_int, _P, C.POINTER(_P), _P]
        L.flsgpu_file_load.restype = C.c_int
        L.flsgpu_file_read_host.argtypes = [C.c_char_p, C.c_int, _P, C.POINTER(_P)]
        L.flsgpu_file_read_host.restype = C.c_int
        L.flsgpu_file_free_host.argtypes = [_P]
        L.flsgpu_file_free_host.restype = None
        L.flsgpu_column_upload.argtypes = [C.c_int, _P, _P, C.POINTER(_P), _P]
        L.flsgpu_column_upload.restype = C.c_int
        L.flsgpu_column_free.argtypes = [_P]
        L.flsgpu_column_free.restype = C.c_int
        L._file_sigs = True
    return L


@dataclass
class FforHostColumn:
    """host::FFORColumn<T> / BPColumn<T> (bases None)."""
    n_values: int
    packed: np.ndarray
    bit_widths: np.ndarray
    vector_offsets: np.ndarray
    bases: np.ndarray | None

    @property
    def n_vecs(self):
        return self.n_values // 1024

    @property
    def width(self):
        return self.packed.dtype.itemsize * 8


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def _bp(n_values, packed, bw, offs):
    return A.BPColumn(n_values, n_values // 1024, _p(packed), _p(bw), _p(offs))


def encoding_of(col) -> int:
    if isinstance(col, FforHostColumn):
        return ENC[("bp_u" if col.bases is None else "ffor_u") + str(col.width)]
    if isinstance(col, AlpHostColumn):
        w = col.dtype.itemsize * 8
        return ENC[("alp_ext_f" if col.offsets_counts is not None else "alp_f") + str(w)]
    if isinstance(col, AlprdHostColumn):
        return ENC[f"alprd_f{col.width}"]
    raise TypeError(f"not a host column: {type(col).__name__}")


def _descriptor(col):
    """ctypes descriptor with host pointers; returns (struct, keepalive)."""
    keep = []

    def c(a, dtype=None):
        if a is None:
            return None
        a = np.ascontiguousarray(a if dtype is None else a.astype(dtype, copy=False))
        keep.append(a)
        return a

    if isinstance(col, FforHostColumn):
        bp = _bp(col.n_values, c(col.packed), c(col.bit_widths, np.uint8),
                 c(col.vector_offsets, np.uint64))
        if col.bases is None:
            return bp, keep
        return A.FFORColumn(bp, _p(c(col.bases, col.packed.dtype))), keep
    if isinstance(col, AlpHostColumn):
        bp = _bp(col.n_values, c(col.packed), c(col.bit_widths), c(col.vector_offsets))
        return A.ALPColumn(A.FFORColumn(bp, _p(c(col.bases))), _p(c(col.factor_indices)),
                           _p(c(col.fraction_indices)), len(col.exceptions),
                           _p(c(col.exceptions_offsets)), _p(c(col.exceptions)),
                           _p(c(col.positions)),
                           None if col.offsets_counts is not None else _p(c(col.counts)),
                           _p(c(col.offsets_counts))), keep
    if isinstance(col, AlprdHostColumn):
        d = (C.c_uint16 * 8)(*([int(x) for x in col.dictionary] + [0] * (8 - len(col.dictionary))))
        return A.ALPRDColumn(col.n_values, col.n_vecs, col.right_bw, col.left_bw,
                             len(col.dictionary), d, _p(c(col.right_packed)),
                             _p(c(col.left_packed)), _p(c(col.exceptions_offsets)),
                             _p(c(col.exceptions)), _p(c(col.positions)), _p(c(col.counts))), keep
    raise TypeError(f"not a host column: {type(col).__name__}")


def _struct_for(encoding: int):
    name = ENC_NAME[encoding]
    if name.startswith("bp_"):
        return A.BPColumn()
    if name.startswith("ffor_"):
        return A.FFORColumn()
    if name.startswith("alprd_"):
        return A.ALPRDColumn()
    return A.ALPColumn()


def write(path: str, col) -> None:
    st, keep = _descriptor(col)
    A.check(_lib().flsgpu_file_write(path.encode(), encoding_of(col), C.byref(st)))
    del keep


def stat(path: str) -> dict:
    info = FileInfo()
    A.check(_lib().flsgpu_file_stat(path.encode(), C.byref(info)))
    return {"encoding": ENC_NAME[info.encoding], "n_values": info.n_values,
            "n_vecs": info.n_vecs, "payload_bytes": info.payload_bytes,
            "file_bytes": info.file_bytes}


def _np_view(addr, count, dtype):
    if count == 0 or not addr:
        return np.zeros(0, dtype)
    buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype, count=count).copy()


def read_host(path: str):
    """Reads a column file back into a host column (copies of the arrays)."""
    enc = ENC[stat(path)["encoding"]]
    st = _struct_for(enc)
    block = _P()
    A.check(_lib().flsgpu_file_read_host(path.encode(), enc, C.byref(st), C.byref(block)))
    try:
        name = ENC_NAME[enc]
        w = int(name[-2:])
        nl = 1024 // w
        if name.startswith("alprd_"):
            nv = st.n_vecs
            cnt = _np_view(st.counts, nv, np.uint16)
            eo = _np_view(st.exceptions_offsets, nv, np.uint64)
            ne = int((eo + cnt).max()) if nv else 0
            return AlprdHostColumn(
                st.n_values, st.right_bit_width, st.left_bit_width,
                np.array(list(st.dictionary)[: st.dictionary_size], np.uint16),
                _np_view(st.right_packed, nv * st.right_bit_width * nl, _UT[w]),
                _np_view(st.left_packed, nv * st.left_bit_width * 64, np.uint16), eo,
                _np_view(st.exceptions, ne, np.uint16), _np_view(st.positions, ne, np.uint16), cnt)
        bp = st.ffor.bp if name.startswith("alp") else (st.bp if name.startswith("ffor") else st)
        nv = bp.n_vecs
        bw = _np_view(bp.bit_widths, nv, np.uint8)
        offs = _np_view(bp.vector_offsets, nv, np.uint64)
        words = int((offs + bw.astype(np.uint64) * np.uint64(nl)).max()) if nv else 0
        packed = _np_view(bp.packed_array, words, _UT[w])
        if name.startswith("bp_"):
            return FforHostColumn(bp.n_values, packed, bw, offs, None)
        if name.startswith("ffor_"):
            return FforHostColumn(bp.n_values, packed, bw, offs, _np_view(st.bases, nv, _UT[w]))
        ne = st.n_exceptions
        ext = name.startswith("alp_ext")
        return AlpHostColumn(
            bp.n_values, packed, bw, offs, _np_view(st.ffor.bases, nv, _UT[w]),
            _np_view(st.factor_indices, nv, np.uint8), _np_view(st.fraction_indices, nv, np.uint8),
            _np_view(st.exceptions_offsets, nv, np.uint64), _np_view(st.exceptions, ne, _FT[w]),
            _np_view(st.positions, ne, np.uint16),
            np.zeros(nv, np.uint16) if ext else _np_view(st.counts, nv, np.uint16),
            _np_view(st.offsets_counts, nv * nl, np.uint16) if ext else None)
    finally:
        _lib().flsgpu_file_free_host(block)


class BlockColumn:
    """A column held in one device block owned by the library (a loaded
    column file or an uploaded host column); freed with the object."""

    def __init__(self, encoding: int, fill):
        self.encoding = encoding
        self.name = ENC_NAME[self.encoding]
        self.struct = _struct_for(self.encoding)
        self._block = _P()
        A.check(fill(self.struct, self._block))
        self.width = int(self.name[-2:])
        bp = (self.struct.ffor.bp if self.name.startswith("alp_") else
              self.struct.bp if self.name.startswith("ffor_") else self.struct)
        self.n_values = int(self.struct.n_values if self.name.startswith("alprd_") else bp.n_values)

    def __del__(self):
        if getattr(self, "_block", None) and self._block.value:
            _lib().flsgpu_column_free(self._block)
            self._block = _P()

    def out_dtype(self):
        import torch
        if self.name.startswith(("bp_", "ffor_")):
            return torch.uint32 if self.width == 32 else torch.uint64
        return torch.float32 if self.width == 32 else torch.float64

    def decompress(self, out=None, variant: int = 0, stream=None):
        import torch
        from .device import _stream_ptr
        if out is None:
            out = torch.empty(self.n_values, dtype=self.out_dtype(), device="cuda")
        L = A.lib()
        s = _stream_ptr(stream)
        n = self.name
        if n.startswith(("bp_", "ffor_")):
            fn = getattr(L, f"flsgpu_decompress_{n}")
            A.check(fn(C.byref(self.struct), out.data_ptr(), variant, s))
        elif n.startswith("alprd_"):
            fn = getattr(L, f"flsgpu_decompress_{n}")
            A.check(fn(C.byref(self.struct), out.data_ptr(), variant, s))
        else:
            patcher = A.PATCHERS["prefetch-all" if n.startswith("alp_ext") else "stateful"]
            fn = getattr(L, f"flsgpu_decompress_{n}")
            A.check(fn(C.byref(self.struct), out.data_ptr(), patcher, variant, s))
        return out

    def query_any(self, magic, stream=None):
        import torch
        from .device import _stream_ptr
        if self.name.startswith("alprd_"):
            raise A.FlsgpuError(A.FLSGPU_EUNSUPPORTED, "no ALP-RD query kernel")
        r = torch.empty(1, dtype=torch.uint8, device="cuda")
        L = A.lib()
        n = self.name
        s = _stream_ptr(stream)
        fn = getattr(L, f"flsgpu_query_any_{n}")
        if n.startswith(("bp_", "ffor_")):
            A.check(fn(C.byref(self.struct), magic, r.data_ptr(), 0, s))
        else:
            patcher = A.PATCHERS["prefetch-all" if n.startswith("alp_ext") else "stateful"]
            A.check(fn(C.byref(self.struct), magic, r.data_ptr(), patcher, 0, s))
        return r


def load(path: str, stream=None) -> BlockColumn:
    """flsgpu_file_load: the column file's payload copied into one device block."""
    from .device import _stream_ptr
    enc = ENC[stat(path)["encoding"]]
    return BlockColumn(enc, lambda st, blk: _lib().flsgpu_file_load(
        path.encode(), enc, C.byref(st), C.byref(blk), _stream_ptr(stream)))


def upload(col, stream=None) -> BlockColumn:
    """flsgpu_column_upload: a host column copied into one device block with
    the reference's over-read padding (host::*Column::copy_to_device)."""
    from .device import _stream_ptr
    enc = encoding_of(col)
    host, keep = _descriptor(col)

    def fill(st, blk):
        rc = _lib().flsgpu_column_upload(enc, C.byref(host), C.byref(st), C.byref(blk),
                                         _stream_ptr(stream))
        del keep[:]
        return rc
    return BlockColumn(enc, fill)
