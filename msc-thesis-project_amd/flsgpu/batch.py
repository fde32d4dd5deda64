"""Batched mixed-encoding decode: one persistent launch (optionally replayed
from a hipGraph) for many columns of FFOR / ALP / G-ALP / ALP-RD encodings."""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi as A
from .alp import DeviceAlpColumn, DeviceAlprdColumn
from .device import DeviceFFORColumn, _ptr, _stream_ptr

ENC = {"bp_u32": 0, "ffor_u32": 1, "bp_u64": 2, "ffor_u64": 3, "alp_f32": 4, "alp_f64": 5,
       "alp_ext_f32": 6, "alp_ext_f64": 7, "alprd_f32": 8, "alprd_f64": 9}


class BatchItem(C.Structure):
    _fields_ = [("encoding", C.c_int), ("column", C.c_void_p), ("out", C.c_void_p),
                ("patcher", C.c_int)]


def _sigs():
    L = A.lib()
    if not getattr(L, "_batch_sigs", False):
        L.flsgpu_batch_plan_create.argtypes = [C.POINTER(BatchItem), C.c_uint32, C.c_int,
                                               C.POINTER(C.c_void_p)]
        L.flsgpu_batch_plan_create.restype = C.c_int
        L.flsgpu_batch_plan_tasks.argtypes = [C.c_void_p]
        L.flsgpu_batch_plan_tasks.restype = C.c_uint64
        for n in ("flsgpu_batch_run", "flsgpu_batch_graph_launch"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_void_p]
            getattr(L, n).restype = C.c_int
        L.flsgpu_batch_plan_destroy.argtypes = [C.c_void_p]
        L.flsgpu_batch_plan_destroy.restype = C.c_int
        L._batch_sigs = True
    return L


def encoding_of(col) -> int:
    if isinstance(col, DeviceFFORColumn):
        kind = ("bp" if col.bases is None else "ffor") + f"_u{col.width}"
    elif isinstance(col, DeviceAlpColumn):
        kind = ("alp_ext" if col.extended else "alp") + f"_f{col.width}"
    elif isinstance(col, DeviceAlprdColumn):
        kind = f"alprd_f{col.width}"
    else:
        raise TypeError(type(col))
    return ENC[kind]


def _descriptor(col):
    if isinstance(col, DeviceFFORColumn):
        return col.bp_struct() if col.bases is None else col.ffor_struct()
    return col.struct()


class BatchPlan:
    """flsgpu_batch_plan over (column, output tensor) pairs."""

    def __init__(self, columns, outputs, variant: int = 0):
        L = _sigs()
        self._descs = [_descriptor(c) for c in columns]
        items = (BatchItem * len(columns))()
        for i, (c, o) in enumerate(zip(columns, outputs)):
            patcher = 0
            if isinstance(c, DeviceAlpColumn):
                patcher = A.PATCHERS["prefetch-all" if c.extended else "stateful"]
            items[i] = BatchItem(encoding_of(c), C.cast(C.pointer(self._descs[i]), C.c_void_p),
                                 _ptr(o), patcher)
        self._items = items
        self._outputs = outputs
        h = C.c_void_p()
        A.check(L.flsgpu_batch_plan_create(items, len(columns), variant, C.byref(h)))
        self.handle = h

    @property
    def n_tasks(self) -> int:
        return int(_sigs().flsgpu_batch_plan_tasks(self.handle))

    def run(self, stream=None):
        A.check(_sigs().flsgpu_batch_run(self.handle, _stream_ptr(stream)))

    def graph_launch(self, stream=None):
        A.check(_sigs().flsgpu_batch_graph_launch(self.handle, _stream_ptr(stream)))

    def close(self):
        if self.handle:
            _sigs().flsgpu_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_each(columns, outputs, stream=None, variant: int = 0):
    """One plain launch per column (the reference's execution model)."""
    from . import alp as AL
    from . import device as D
    for c, o in zip(columns, outputs):
        if isinstance(c, D.DeviceFFORColumn):
            D.decompress(c, o, variant=variant, stream=stream)
        elif isinstance(c, AL.DeviceAlpColumn):
            AL.decompress(c, o, variant=variant, stream=stream)
        else:
            AL.alprd_decompress(c, o, variant=variant, stream=stream)


def output_for(col, device="cuda") -> torch.Tensor:
    if isinstance(col, DeviceFFORColumn):
        dt = torch.uint32 if col.width == 32 else torch.uint64
    else:
        dt = torch.float32 if col.width == 32 else torch.float64
    return torch.empty(col.n_values, dtype=dt, device=device)
