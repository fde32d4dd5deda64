"""ctypes face of the CPU oracle (oracle/_build/liboracle.so) and of the
compiled reference codec (oracle/_ref/libflsref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker or the timed baseline,
never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libflsref.so")

_lib = None
_ref = None

_NP = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}
_CT = {8: C.c_uint8, 16: C.c_uint16, 32: C.c_uint32, 64: C.c_uint64}


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
        _lib = C.CDLL(ORACLE_SO)
        for w in (8, 16, 32, 64):
            for n in ("pack", "unpack"):
                f = getattr(_lib, f"orc_fls_{n}_u{w}")
                f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint]
                f.restype = None
            for n in ("ffor", "unffor"):
                f = getattr(_lib, f"orc_fls_{n}_u{w}")
                f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, _CT[w]]
                f.restype = None
        for sfx, ut in (("f32", C.c_uint32), ("f64", C.c_uint64)):
            f = getattr(_lib, f"orc_alp_decode_vec_{sfx}")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, ut, C.c_uint, C.c_uint]
            f.restype = None
            f = getattr(_lib, f"orc_alp_encode_value_{sfx}")
            f.argtypes = [C.c_float if sfx == "f32" else C.c_double, C.c_uint, C.c_uint, C.c_int]
            f.restype = C.c_int32 if sfx == "f32" else C.c_int64
            f = getattr(_lib, f"orc_alp_sample_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]
            f.restype = C.c_size_t
            f = getattr(_lib, f"orc_alp_top_k_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
            f.restype = C.c_int
            f = getattr(_lib, f"orc_alp_search_one_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
            f.restype = C.c_uint64
            f = getattr(_lib, f"orc_alp_estimate_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_uint, C.c_uint]
            f.restype = C.c_uint64
            f = getattr(_lib, f"orc_alp_best_combination_{sfx}")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = None
            f = getattr(_lib, f"orc_alp_encode_vec_{sfx}")
            f.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = C.c_uint
            f = getattr(_lib, f"orc_alp_analyze_ffor_{sfx}")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = None
            f = getattr(_lib, f"orc_alp_column_encode_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t] + [C.c_void_p] * 11
            f.restype = C.c_int
            f = getattr(_lib, f"orc_alp_column_decode_{sfx}")
            f.argtypes = [C.c_size_t] + [C.c_void_p] * 12
            f.restype = None
            f = getattr(_lib, f"orc_galp_convert_{sfx}")
            f.argtypes = [C.c_size_t] + [C.c_void_p] * 6
            f.restype = None
            f = getattr(_lib, f"orc_alprd_find_dictionary_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = C.c_uint
            f = getattr(_lib, f"orc_alprd_column_encode_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_uint, C.c_void_p, C.c_uint, C.c_uint] + [C.c_void_p] * 6
            f.restype = None
            f = getattr(_lib, f"orc_alprd_column_decode_{sfx}")
            f.argtypes = [C.c_size_t, C.c_uint, C.c_void_p, C.c_uint] + [C.c_void_p] * 7
            f.restype = None
        for w in (32, 64):
            f = getattr(_lib, f"orc_ffor_column_decode_u{w}")
            f.argtypes = [C.c_size_t] + [C.c_void_p] * 5
            f.restype = None
        _lib.orc_ffor_column_decode_u32_mt.argtypes = [C.c_size_t] + [C.c_void_p] * 5 + [C.c_int]
        _lib.orc_ffor_column_decode_u32_mt.restype = None
        _lib.orc_alp_column_decode_f64_mt.argtypes = [C.c_size_t] + [C.c_void_p] * 11 + [C.c_int]
        _lib.orc_alp_column_decode_f64_mt.restype = None
    return _lib


def ref():
    """The compiled reference codec, or None when it was not built here."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        _ref = C.CDLL(REF_SO)
        for w in (8, 16, 32, 64):
            for n in ("pack", "unpack"):
                f = getattr(_ref, f"ref_fls_{n}_u{w}")
                f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint]
                f.restype = None
            for n in ("ffor", "unffor"):
                f = getattr(_ref, f"ref_fls_{n}_u{w}")
                f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, _CT[w]]
                f.restype = None
        for w in (32, 64):
            f = getattr(_ref, f"ref_ffor_column_decode_u{w}")
            f.argtypes = [C.c_size_t] + [C.c_void_p] * 5
            f.restype = None
        for sfx, ut in (("f32", C.c_uint32), ("f64", C.c_uint64)):
            f = getattr(_ref, f"ref_falp_{sfx}")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, ut, C.c_uint, C.c_uint]
            f.restype = None
            f = getattr(_ref, f"ref_patch_{sfx}")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint16]
            f.restype = None
            f = getattr(_ref, f"ref_alp_sample_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p]
            f.restype = C.c_size_t
            f = getattr(_ref, f"ref_alp_encode_value_{sfx}")
            f.argtypes = [C.c_float if sfx == "f32" else C.c_double, C.c_uint, C.c_uint]
            f.restype = C.c_int32 if sfx == "f32" else C.c_int64
            f = getattr(_ref, f"ref_alp_top_k_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
            f.restype = C.c_int
            f = getattr(_ref, f"ref_alp_best_combination_{sfx}")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = None
            f = getattr(_ref, f"ref_alp_encode_vec_{sfx}")
            f.argtypes = [C.c_void_p, C.c_uint, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = C.c_uint
            f = getattr(_ref, f"ref_alp_analyze_ffor_{sfx}")
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = None
            f = getattr(_ref, f"ref_alp_column_encode_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t] + [C.c_void_p] * 10
            f.restype = C.c_int
            f = getattr(_ref, f"ref_alprd_find_dictionary_{sfx}")
            f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
            f.restype = C.c_uint
            f = getattr(_ref, f"ref_alp_column_decode_{sfx}")
            f.argtypes = [C.c_size_t, C.c_size_t] + [C.c_void_p] * 11
            f.restype = None
        _ref.ref_alprd_decode_vec_f64.argtypes = [C.c_void_p] * 5 + [C.c_uint16, C.c_void_p, C.c_uint, C.c_uint]
        _ref.ref_alprd_decode_vec_f64.restype = None
    return _ref


# ---------------------------------------------------------------------------
# FastLanes, one vector
# ---------------------------------------------------------------------------
def _vec(which, w, fn, *a):
    getattr(which, fn)(*a)


def pack(values: np.ndarray, bw: int, use_ref=False) -> np.ndarray:
    w = values.dtype.itemsize * 8
    out = np.zeros(bw * (1024 // w), dtype=values.dtype)
    L = ref() if use_ref else lib()
    getattr(L, ("ref" if use_ref else "orc") + f"_fls_pack_u{w}")(_p(values), _p(out), bw)
    return out


def unpack(packed: np.ndarray, bw: int, use_ref=False) -> np.ndarray:
    w = packed.dtype.itemsize * 8
    out = np.zeros(1024, dtype=packed.dtype)
    src = np.ascontiguousarray(packed)
    L = ref() if use_ref else lib()
    getattr(L, ("ref" if use_ref else "orc") + f"_fls_unpack_u{w}")(_p(src), _p(out), bw)
    return out


def ffor(values: np.ndarray, bw: int, base: int, use_ref=False) -> np.ndarray:
    w = values.dtype.itemsize * 8
    out = np.zeros(bw * (1024 // w), dtype=values.dtype)
    L = ref() if use_ref else lib()
    getattr(L, ("ref" if use_ref else "orc") + f"_fls_ffor_u{w}")(_p(values), _p(out), bw, int(base))
    return out


def unffor(packed: np.ndarray, bw: int, base: int, use_ref=False) -> np.ndarray:
    w = packed.dtype.itemsize * 8
    out = np.zeros(1024, dtype=packed.dtype)
    src = np.ascontiguousarray(packed)
    L = ref() if use_ref else lib()
    getattr(L, ("ref" if use_ref else "orc") + f"_fls_unffor_u{w}")(_p(src), _p(out), bw, int(base))
    return out


# ---------------------------------------------------------------------------
# Column containers (host, numpy): field names follow src/flsgpu/structs.cuh
# ---------------------------------------------------------------------------
@dataclass
class FFORColumn:
    n_values: int
    packed: np.ndarray          # UINT words
    bit_widths: np.ndarray      # uint8 [n_vecs]
    vector_offsets: np.ndarray  # uint64 [n_vecs], in words
    bases: np.ndarray | None    # UINT [n_vecs]; None for plain bit-packing

    @property
    def n_vecs(self):
        return len(self.bit_widths)


@dataclass
class ALPColumn:
    n_values: int
    ffor: FFORColumn
    factor_indices: np.ndarray
    fraction_indices: np.ndarray
    exceptions_offsets: np.ndarray
    exceptions: np.ndarray
    positions: np.ndarray
    counts: np.ndarray | None = None          # standard layout
    offsets_counts: np.ndarray | None = None  # G-ALP layout
    compressed_bytes_alp: int = 0
    compressed_bytes_galp: int = 0

    @property
    def n_vecs(self):
        return self.ffor.n_vecs


@dataclass
class ALPRDColumn:
    n_values: int
    right_bw: int
    left_bw: int
    dictionary: np.ndarray      # uint16 [dict_size]
    right_packed: np.ndarray    # UINT words, n_vecs * right_bw * N_LANES
    left_packed: np.ndarray     # uint16 words, n_vecs * left_bw * 64
    exceptions_offsets: np.ndarray
    exceptions: np.ndarray      # uint16 original left parts
    positions: np.ndarray
    counts: np.ndarray

    @property
    def n_vecs(self):
        return self.n_values // 1024


def ffor_column_decode(col: FFORColumn, use_ref=False) -> np.ndarray:
    w = col.packed.dtype.itemsize * 8
    out = np.zeros(col.n_vecs * 1024, dtype=col.packed.dtype)
    bases = col.bases if col.bases is not None else np.zeros(col.n_vecs, dtype=col.packed.dtype)
    L = ref() if use_ref else lib()
    name = ("ref" if use_ref else "orc") + f"_ffor_column_decode_u{w}"
    getattr(L, name)(col.n_vecs, _p(col.packed), _p(col.bit_widths), _p(col.vector_offsets),
                     _p(bases), _p(out))
    return out


def alp_decode_vec(packed, bw, base, fac, exp, dtype, use_ref=False):
    out = np.zeros(1024, dtype=dtype)
    sfx = "f32" if np.dtype(dtype) == np.float32 else "f64"
    src = np.ascontiguousarray(packed)
    if use_ref:
        getattr(ref(), f"ref_falp_{sfx}")(_p(src), _p(out), bw, int(base), fac, exp)
    else:
        getattr(lib(), f"orc_alp_decode_vec_{sfx}")(_p(src), _p(out), bw, int(base), fac, exp)
    return out


def ref_alp_column_decode_range(col, v0: int, v1: int, out: np.ndarray) -> None:
    """The compiled reference's CPU decode (falp + patch_exceptions per
    vector, src/alp/alp-bindings.cu:208-228) of vectors [v0, v1) of a
    standard-layout ALP column into out[v0 * 1024 : v1 * 1024]. `col` is any
    object with the ALPColumn array fields at top level (packed, bit_widths,
    vector_offsets, bases, factor_indices, fraction_indices, exceptions,
    positions, exceptions_offsets, counts)."""
    sfx = "f32" if out.dtype == np.float32 else "f64"
    getattr(ref(), f"ref_alp_column_decode_{sfx}")(
        v0, v1, _p(col.packed), _p(col.bit_widths), _p(col.vector_offsets), _p(col.bases),
        _p(col.factor_indices), _p(col.fraction_indices), _p(col.exceptions), _p(col.positions),
        _p(col.exceptions_offsets), _p(col.counts), _p(out))


def alp_column_decode(col: ALPColumn) -> np.ndarray:
    dtype = col.exceptions.dtype
    sfx = "f32" if dtype == np.float32 else "f64"
    out = np.zeros(col.n_vecs * 1024, dtype=dtype)
    getattr(lib(), f"orc_alp_column_decode_{sfx}")(
        col.n_vecs, _p(col.ffor.packed), _p(col.ffor.bit_widths), _p(col.ffor.vector_offsets),
        _p(col.ffor.bases), _p(col.factor_indices), _p(col.fraction_indices),
        _p(col.exceptions_offsets), _p(col.exceptions), _p(col.positions), _p(col.counts),
        _p(col.offsets_counts), _p(out))
    return out


def galp_convert(col: ALPColumn) -> ALPColumn:
    """Standard -> G-ALP lane-divided layout (structs.cuh:231-331)."""
    dtype = col.exceptions.dtype
    sfx = "f32" if dtype == np.float32 else "f64"
    n_lanes = 32 if dtype == np.float32 else 16
    exc = np.zeros(max(len(col.exceptions), 1), dtype=dtype)
    pos = np.zeros(max(len(col.positions), 1), dtype=np.uint16)
    oc = np.zeros(col.n_vecs * n_lanes, dtype=np.uint16)
    getattr(lib(), f"orc_galp_convert_{sfx}")(col.n_vecs, _p(col.exceptions), _p(col.positions),
                                              _p(col.counts), _p(exc), _p(pos), _p(oc))
    n = len(col.exceptions)
    return ALPColumn(col.n_values, col.ffor, col.factor_indices, col.fraction_indices,
                     col.exceptions_offsets, exc[:n], pos[:n], None, oc,
                     col.compressed_bytes_alp, col.compressed_bytes_galp)


def alp_column_encode(values: np.ndarray, use_ref=False) -> ALPColumn | None:
    """alp::encode (alp-bindings.cu:83-206). None when the column samples as ALP_RD."""
    dtype = values.dtype
    sfx = "f32" if dtype == np.float32 else "f64"
    ut = np.uint32 if dtype == np.float32 else np.uint64
    n = len(values)
    assert n % 1024 == 0
    nv = n // 1024
    packed = np.zeros(n + 64, dtype=ut)
    bw = np.zeros(nv, dtype=np.uint8)
    vo = np.zeros(nv, dtype=np.uint64)
    bases = np.zeros(nv, dtype=ut)
    fi = np.zeros(nv, dtype=np.uint8)
    ri = np.zeros(nv, dtype=np.uint8)
    eo = np.zeros(nv, dtype=np.uint64)
    exc = np.zeros(n + 1, dtype=dtype)
    pos = np.zeros(n + 1, dtype=np.uint16)
    cnt = np.zeros(nv, dtype=np.uint16)
    sizes = np.zeros(4, dtype=np.uint64)
    values = np.ascontiguousarray(values)
    if use_ref:
        rc = getattr(ref(), f"ref_alp_column_encode_{sfx}")(
            _p(values), nv, _p(packed), _p(bw), _p(vo), _p(bases), _p(fi), _p(ri), _p(eo),
            _p(exc), _p(pos), _p(cnt))
        n_words = int((bw.astype(np.uint64) * (1024 // (8 * values.itemsize))).sum())
        n_exc = int(cnt.astype(np.uint64).sum())
    else:
        rc = getattr(lib(), f"orc_alp_column_encode_{sfx}")(
            _p(values), nv, _p(packed), _p(bw), _p(vo), _p(bases), _p(fi), _p(ri), _p(eo),
            _p(exc), _p(pos), _p(cnt), _p(sizes))
        n_words, n_exc = int(sizes[0]), int(sizes[1])
    if rc != 0:
        return None
    ffor_col = FFORColumn(n, packed[:n_words].copy(), bw, vo, bases)
    return ALPColumn(n, ffor_col, fi, ri, eo, exc[:n_exc].copy(), pos[:n_exc].copy(), cnt, None,
                     int(sizes[2]), int(sizes[3]))


def alp_sample(values: np.ndarray, use_ref=False) -> np.ndarray:
    sfx = "f32" if values.dtype == np.float32 else "f64"
    out = np.zeros(min(len(values), 102400) + 64, dtype=values.dtype)
    values = np.ascontiguousarray(values)
    if use_ref:
        n = getattr(ref(), f"ref_alp_sample_{sfx}")(_p(values), 0, len(values), _p(out))
    else:
        n = getattr(lib(), f"orc_alp_sample_{sfx}")(_p(values), 0, len(values), _p(out))
    return out[:n].copy()


def alp_top_k(values_or_sample: np.ndarray, use_ref=False):
    """Oracle: top-k on a sample. Reference: encoder<T>::init on a column."""
    sfx = "f32" if values_or_sample.dtype == np.float32 else "f64"
    e = np.zeros(8, dtype=np.uint8)
    f = np.zeros(8, dtype=np.uint8)
    a = np.ascontiguousarray(values_or_sample)
    if use_ref:
        k = getattr(ref(), f"ref_alp_top_k_{sfx}")(_p(a), len(a), _p(e), _p(f))
    else:
        k = getattr(lib(), f"orc_alp_top_k_{sfx}")(_p(a), len(a), _p(e), _p(f))
    return [(int(e[i]), int(f[i])) for i in range(k)]


def alp_search_one(sample32: np.ndarray):
    sfx = "f32" if sample32.dtype == np.float32 else "f64"
    e = np.zeros(1, dtype=np.uint8)
    f = np.zeros(1, dtype=np.uint8)
    a = np.ascontiguousarray(sample32)
    est = getattr(lib(), f"orc_alp_search_one_{sfx}")(_p(a), len(a), _p(e), _p(f))
    return int(e[0]), int(f[0]), int(est)


def alp_estimate(sample32: np.ndarray, exp: int, fac: int):
    """Estimated bits of (exp, fac) on the samples; None if skipped."""
    sfx = "f32" if sample32.dtype == np.float32 else "f64"
    a = np.ascontiguousarray(sample32)
    v = getattr(lib(), f"orc_alp_estimate_{sfx}")(_p(a), len(a), exp, fac)
    return None if v == 2 ** 64 - 1 else int(v)


def alp_best_combination(combos, vec: np.ndarray, use_ref=False):
    sfx = "f32" if vec.dtype == np.float32 else "f64"
    e = np.array([c[0] for c in combos], dtype=np.uint8)
    f = np.array([c[1] for c in combos], dtype=np.uint8)
    oe = np.zeros(1, dtype=np.uint8)
    of = np.zeros(1, dtype=np.uint8)
    vec = np.ascontiguousarray(vec)
    L = ref() if use_ref else lib()
    getattr(L, ("ref" if use_ref else "orc") + f"_alp_best_combination_{sfx}")(
        _p(e), _p(f), len(combos), _p(vec), _p(oe), _p(of))
    return int(oe[0]), int(of[0])


def alp_encode_vec(vec: np.ndarray, fac: int, exp: int, use_ref=False):
    sfx = "f32" if vec.dtype == np.float32 else "f64"
    st = np.int32 if vec.dtype == np.float32 else np.int64
    enc = np.zeros(1024, dtype=st)
    exc = np.zeros(1024, dtype=vec.dtype)
    pos = np.zeros(1024, dtype=np.uint16)
    vec = np.ascontiguousarray(vec)
    L = ref() if use_ref else lib()
    c = getattr(L, ("ref" if use_ref else "orc") + f"_alp_encode_vec_{sfx}")(
        _p(vec), fac, exp, _p(enc), _p(exc), _p(pos))
    return enc, exc[:c].copy(), pos[:c].copy()


def alp_analyze_ffor(enc: np.ndarray, use_ref=False):
    sfx = "f32" if enc.dtype == np.int32 else "f64"
    bw = np.zeros(1, dtype=np.uint8)
    base = np.zeros(1, dtype=enc.dtype)
    L = ref() if use_ref else lib()
    getattr(L, ("ref" if use_ref else "orc") + f"_alp_analyze_ffor_{sfx}")(_p(enc), _p(bw), _p(base))
    return int(bw[0]), int(base[0])


def alp_encode_value(v, fac, exp, use_ref=False):
    sfx = "f32" if isinstance(v, np.float32) else "f64"
    if use_ref:
        return getattr(ref(), f"ref_alp_encode_value_{sfx}")(v, fac, exp)
    return getattr(lib(), f"orc_alp_encode_value_{sfx}")(v, fac, exp, 1)


def alprd_find_dictionary(sample: np.ndarray, use_ref=False):
    sfx = "f32" if sample.dtype == np.float32 else "f64"
    d = np.zeros(8, dtype=np.uint16)
    ds = C.c_uint(0)
    lbw = C.c_uint(0)
    s = np.ascontiguousarray(sample)
    L = ref() if use_ref else lib()
    rbw = getattr(L, ("ref" if use_ref else "orc") + f"_alprd_find_dictionary_{sfx}")(
        _p(s), len(s), _p(d), C.byref(ds), C.byref(lbw))
    return int(rbw), d[: ds.value].copy(), int(lbw.value)


def alprd_column_encode(values: np.ndarray) -> ALPRDColumn:
    sfx = "f32" if values.dtype == np.float32 else "f64"
    w = values.itemsize * 8
    ut = np.uint32 if w == 32 else np.uint64
    n = len(values)
    nv = n // 1024
    sample = alp_sample(values)
    rbw, d, lbw = alprd_find_dictionary(sample)
    rp = np.zeros(max(nv * rbw * (1024 // w), 1), dtype=ut)
    lp = np.zeros(max(nv * lbw * 64, 1), dtype=np.uint16)
    eo = np.zeros(nv, dtype=np.uint64)
    exc = np.zeros(n + 1, dtype=np.uint16)
    pos = np.zeros(n + 1, dtype=np.uint16)
    cnt = np.zeros(nv, dtype=np.uint16)
    dd = np.zeros(8, dtype=np.uint16)
    dd[: len(d)] = d
    values = np.ascontiguousarray(values)
    getattr(lib(), f"orc_alprd_column_encode_{sfx}")(_p(values), nv, rbw, _p(dd), len(d), lbw,
                                                     _p(rp), _p(lp), _p(eo), _p(exc), _p(pos), _p(cnt))
    ne = int(cnt.astype(np.uint64).sum())
    return ALPRDColumn(n, rbw, lbw, dd[: len(d)].copy(), rp, lp, eo, exc[:ne].copy(), pos[:ne].copy(), cnt)


def alprd_column_decode(col: ALPRDColumn) -> np.ndarray:
    w = col.right_packed.dtype.itemsize * 8
    sfx = "f32" if w == 32 else "f64"
    out = np.zeros(col.n_vecs * 1024, dtype=np.float32 if w == 32 else np.float64)
    dd = np.zeros(8, dtype=np.uint16)
    dd[: len(col.dictionary)] = col.dictionary
    exc = col.exceptions if len(col.exceptions) else np.zeros(1, np.uint16)
    pos = col.positions if len(col.positions) else np.zeros(1, np.uint16)
    getattr(lib(), f"orc_alprd_column_decode_{sfx}")(
        col.n_vecs, col.right_bw, _p(dd), col.left_bw, _p(col.right_packed), _p(col.left_packed),
        _p(col.exceptions_offsets), _p(exc), _p(pos), _p(col.counts), _p(out))
    return out
