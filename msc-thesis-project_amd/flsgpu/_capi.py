"""ctypes binding of the C ABI declared in include/flsgpu.h.

The shared library is built in-tree (msc-thesis-project_amd/lib/libflsgpu_hip.so).
There is no fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libflsgpu_hip.so")
HEADER_PATH = os.path.join(REPO_ROOT, "include", "flsgpu.h")

FLSGPU_OK = 0
FLSGPU_EINVAL = -1
FLSGPU_EUNSUPPORTED = -2
FLSGPU_EHIP = -3

VARIANT_AUTO = 0
VARIANT_LANE = 1
VARIANT_LANE_NT = 2
VARIANT_HALF_LANE = 3
VARIANT_HALF_LANE_NT = 4
VARIANT_QUARTER_LANE = 5
VARIANT_QUARTER_LANE_NT = 6
VARIANT_HALF_LANE_STAGED = 7
VARIANT_QUARTER_LANE_STAGED = 8
VARIANT_EIGHTH_LANE_STAGED = 9
VARIANTS = {
    "lane": VARIANT_LANE,
    "lane-nt": VARIANT_LANE_NT,
    "half-lane": VARIANT_HALF_LANE,
    "half-lane-nt": VARIANT_HALF_LANE_NT,
    "quarter-lane": VARIANT_QUARTER_LANE,
    "quarter-lane-nt": VARIANT_QUARTER_LANE_NT,
    "half-lane-staged": VARIANT_HALF_LANE_STAGED,
    "quarter-lane-staged": VARIANT_QUARTER_LANE_STAGED,
    "eighth-lane-staged": VARIANT_EIGHTH_LANE_STAGED,
}

PATCHERS = {
    "none": 0, "dummy": 1, "stateless": 2, "stateful": 3, "naive": 4,
    "naive-branchless": 5, "prefetch-position": 6, "prefetch-all": 7,
    "prefetch-all-branchless": 8,
}

PRED_RANGE = 0
PRED_OFFSET_LOWER_HALF = 1
PRED_ALL = 2


class BPColumn(C.Structure):
    _fields_ = [
        ("n_values", C.c_uint64),
        ("n_vecs", C.c_uint64),
        ("packed_array", C.c_void_p),
        ("bit_widths", C.c_void_p),
        ("vector_offsets", C.c_void_p),
    ]


class FFORColumn(C.Structure):
    _fields_ = [("bp", BPColumn), ("bases", C.c_void_p)]


class ALPColumn(C.Structure):
    _fields_ = [
        ("ffor", FFORColumn),
        ("factor_indices", C.c_void_p),
        ("fraction_indices", C.c_void_p),
        ("n_exceptions", C.c_uint64),
        ("exceptions_offsets", C.c_void_p),
        ("exceptions", C.c_void_p),
        ("positions", C.c_void_p),
        ("counts", C.c_void_p),
        ("offsets_counts", C.c_void_p),
    ]


class ALPRDColumn(C.Structure):
    _fields_ = [
        ("n_values", C.c_uint64),
        ("n_vecs", C.c_uint64),
        ("right_bit_width", C.c_uint32),
        ("left_bit_width", C.c_uint32),
        ("dictionary_size", C.c_uint32),
        ("dictionary", C.c_uint16 * 8),
        ("right_packed", C.c_void_p),
        ("left_packed", C.c_void_p),
        ("exceptions_offsets", C.c_void_p),
        ("exceptions", C.c_void_p),
        ("positions", C.c_void_p),
        ("counts", C.c_void_p),
    ]


class FlsgpuError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"flsgpu error {code}: {message}")
        self.code = code


_lib = None

_P = C.c_void_p
_SIGS = {
    "flsgpu_last_error": ([], C.c_char_p),
    "flsgpu_init": ([C.c_int], C.c_int),
    "flsgpu_set_tuning": ([C.c_uint32, C.c_uint32, C.c_int], C.c_int),
    "flsgpu_get_tuning": ([C.c_uint32, C.c_uint32], C.c_int),
    "flsgpu_upload": ([C.POINTER(C.c_void_p), _P, C.c_size_t, C.c_size_t, _P], C.c_int),
    "flsgpu_download": ([_P, _P, C.c_size_t, _P], C.c_int),
    "flsgpu_free": ([_P], C.c_int),
    "flsgpu_decompress_bp_u32": ([C.POINTER(BPColumn), _P, C.c_int, _P], C.c_int),
    "flsgpu_decompress_bp_u64": ([C.POINTER(BPColumn), _P, C.c_int, _P], C.c_int),
    "flsgpu_decompress_ffor_u32": ([C.POINTER(FFORColumn), _P, C.c_int, _P], C.c_int),
    "flsgpu_decompress_ffor_u64": ([C.POINTER(FFORColumn), _P, C.c_int, _P], C.c_int),
    "flsgpu_decompress_alp_f32": ([C.POINTER(ALPColumn), _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_decompress_alp_f64": ([C.POINTER(ALPColumn), _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_decompress_alp_ext_f32": ([C.POINTER(ALPColumn), _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_decompress_alp_ext_f64": ([C.POINTER(ALPColumn), _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_decompress_alprd_f32": ([C.POINTER(ALPRDColumn), _P, C.c_int, _P], C.c_int),
    "flsgpu_decompress_alprd_f64": ([C.POINTER(ALPRDColumn), _P, C.c_int, _P], C.c_int),
    "flsgpu_query_any_bp_u32": ([C.POINTER(BPColumn), C.c_uint32, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_any_bp_u64": ([C.POINTER(BPColumn), C.c_uint64, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_any_ffor_u32": ([C.POINTER(FFORColumn), C.c_uint32, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_any_ffor_u64": ([C.POINTER(FFORColumn), C.c_uint64, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_any_alp_f32": ([C.POINTER(ALPColumn), C.c_float, _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_query_any_alp_f64": ([C.POINTER(ALPColumn), C.c_double, _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_query_any_alp_ext_f32": ([C.POINTER(ALPColumn), C.c_float, _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_query_any_alp_ext_f64": ([C.POINTER(ALPColumn), C.c_double, _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_synchronize": ([_P], C.c_int),
    "flsgpu_release_stream": ([_P], C.c_int),
    "flsgpu_query_filter_sum_bp_u32": (
        [C.POINTER(BPColumn), C.c_int, C.c_uint64, C.c_uint64, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_filter_sum_bp_u64": (
        [C.POINTER(BPColumn), C.c_int, C.c_uint64, C.c_uint64, _P, C.c_int, _P], C.c_int),
    "flsgpu_compute_bp_u32": ([C.POINTER(BPColumn), C.c_uint32, _P, C.c_int, _P], C.c_int),
    "flsgpu_compute_bp_u64": ([C.POINTER(BPColumn), C.c_uint32, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_filter_sum_ffor_u32": (
        [C.POINTER(FFORColumn), C.c_int, C.c_uint64, C.c_uint64, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_filter_sum_ffor_u64": (
        [C.POINTER(FFORColumn), C.c_int, C.c_uint64, C.c_uint64, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_filter_sum_ffor_u32_noalloc": (
        [C.POINTER(FFORColumn), C.c_int, C.c_uint64, C.c_uint64, _P, _P], C.c_int),
    "flsgpu_compute_ffor_u32": ([C.POINTER(FFORColumn), C.c_uint32, _P, C.c_int, _P], C.c_int),
    "flsgpu_compute_ffor_u64": ([C.POINTER(FFORColumn), C.c_uint32, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_multi_column_ffor_u32": (
        [C.POINTER(FFORColumn), C.c_uint32, C.c_uint32, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_multi_column_ffor_u64": (
        [C.POINTER(FFORColumn), C.c_uint32, C.c_uint64, _P, C.c_int, _P], C.c_int),
    "flsgpu_query_multi_column_alp_f32": (
        [C.POINTER(ALPColumn), C.c_uint32, C.c_float, _P, C.c_int, C.c_int, _P], C.c_int),
    "flsgpu_query_multi_column_alp_f64": (
        [C.POINTER(ALPColumn), C.c_uint32, C.c_double, _P, C.c_int, C.c_int, _P], C.c_int),
}


def header_symbols(include_dir: str = os.path.dirname(HEADER_PATH)) -> list[str]:
    """Every function name declared in the C-ABI headers (include/*.h)."""
    names = set()
    for f in sorted(os.listdir(include_dir)):
        if f.endswith(".h"):
            text = open(os.path.join(include_dir, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(flsgpu_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def lib() -> C.CDLL:
    """Load the in-tree HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C msc-thesis-project_amd -j8` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != FLSGPU_OK:
        msg = lib().flsgpu_last_error()
        raise FlsgpuError(rc, msg.decode() if msg else "")
