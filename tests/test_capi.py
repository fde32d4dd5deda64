"""C-ABI library checks that need no GPU: the library loads, exports every
symbol include/flsgpu.h declares, and rejects bad arguments with status codes
before touching the device (the reference exits or throws instead:
src/flsgpu/host-utils.cuh:12-20, generate-kernel-bindings.py:207-209)."""
import ctypes as C

import pytest

from flsgpu import _capi as A


def test_library_exports_every_header_symbol():
    L = A.lib()
    names = A.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared in include/flsgpu.h but not exported: {missing}"


def _bad_bp(n_values=1000, n_vecs=1):
    return A.BPColumn(n_values, n_vecs, 0x1000, 0x2000, 0x3000)


def test_rejects_ragged_n_values():
    L = A.lib()
    bp = _bad_bp(1000, 1)  # not n_vecs * 1024
    rc = L.flsgpu_decompress_bp_u32(C.byref(bp), C.c_void_p(0x4000), 0, None)
    assert rc == A.FLSGPU_EINVAL
    assert b"n_values" in L.flsgpu_last_error()


def test_rejects_null_pointers_and_bad_variant():
    L = A.lib()
    bp = A.BPColumn(1024, 1, None, None, None)
    assert L.flsgpu_decompress_bp_u64(C.byref(bp), C.c_void_p(0x4000), 0, None) == A.FLSGPU_EINVAL
    ok = A.BPColumn(1024, 1, 0x1000, 0x2000, 0x3000)
    assert L.flsgpu_decompress_bp_u32(C.byref(ok), None, 0, None) == A.FLSGPU_EINVAL
    assert L.flsgpu_decompress_bp_u32(C.byref(ok), C.c_void_p(0x4000), 99, None) == A.FLSGPU_EUNSUPPORTED
    f = A.FFORColumn(ok, None)
    assert L.flsgpu_decompress_ffor_u32(C.byref(f), C.c_void_p(0x4000), 0, None) == A.FLSGPU_EINVAL
    assert L.flsgpu_query_any_bp_u32(C.byref(ok), 1, None, 0, None) == A.FLSGPU_EINVAL
    f = A.FFORColumn(ok, 0x5000)
    assert L.flsgpu_query_filter_sum_ffor_u32(C.byref(f), 7, 0, 1, C.c_void_p(0x6000), 0, None) in (
        A.FLSGPU_EINVAL,)
    arr = (A.FFORColumn * 11)()
    assert L.flsgpu_query_multi_column_ffor_u32(arr, 11, 0, C.c_void_p(0x6000), 0, None) == A.FLSGPU_EINVAL
    assert L.flsgpu_query_multi_column_ffor_u32(arr, 0, 0, C.c_void_p(0x6000), 0, None) == A.FLSGPU_EINVAL


def test_tuning_table_roundtrip():
    L = A.lib()
    FFOR_U32, FFOR_U64 = 1, 3
    assert L.flsgpu_get_tuning(FFOR_U32, 255) == A.VARIANT_HALF_LANE_NT
    assert L.flsgpu_get_tuning(FFOR_U64, 255) == A.VARIANT_QUARTER_LANE_NT
    assert L.flsgpu_set_tuning(FFOR_U32, 13, A.VARIANT_HALF_LANE) == 0
    assert L.flsgpu_get_tuning(FFOR_U32, 13) == A.VARIANT_HALF_LANE
    assert L.flsgpu_set_tuning(FFOR_U32, 33, A.VARIANT_LANE) == A.FLSGPU_EINVAL
    assert L.flsgpu_set_tuning(FFOR_U64, 64, A.VARIANT_LANE) == 0
    assert L.flsgpu_set_tuning(10, 3, A.VARIANT_LANE) == A.FLSGPU_EINVAL
    assert L.flsgpu_set_tuning(FFOR_U64, 3, 0) == A.FLSGPU_EINVAL
    assert L.flsgpu_set_tuning(FFOR_U32, 13, A.VARIANT_HALF_LANE_NT) == 0
    assert L.flsgpu_set_tuning(FFOR_U64, 64, A.VARIANT_QUARTER_LANE_NT) == 0
