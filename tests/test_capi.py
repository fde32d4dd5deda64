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


def test_filter_sum_and_compute_on_bp_columns_check_arguments():
    L = A.lib()
    bad = A.BPColumn(1000, 1, 0x1000, 0x2000, 0x3000)
    assert L.flsgpu_query_filter_sum_bp_u32(C.byref(bad), 2, 0, 0, C.c_void_p(0x6000), 0,
                                            None) == A.FLSGPU_EINVAL
    ok = A.BPColumn(1024, 1, 0x1000, 0x2000, 0x3000)
    assert L.flsgpu_query_filter_sum_bp_u64(C.byref(ok), 2, 0, 0, None, 0, None) == A.FLSGPU_EINVAL
    assert L.flsgpu_query_filter_sum_bp_u32(C.byref(ok), 9, 0, 0, C.c_void_p(0x6000), 0,
                                            None) == A.FLSGPU_EINVAL
    assert L.flsgpu_compute_bp_u32(C.byref(ok), 1, None, 0, None) == A.FLSGPU_EINVAL
    nul = A.BPColumn(1024, 1, None, 0x2000, 0x3000)
    assert L.flsgpu_compute_bp_u64(C.byref(nul), 1, C.c_void_p(0x6000), 0, None) == A.FLSGPU_EINVAL


def _plan(item_enc, col, patcher=0):
    from flsgpu import batch as B
    L = B._sigs()
    items = (B.BatchItem * 1)(B.BatchItem(item_enc, C.cast(C.pointer(col), C.c_void_p),
                                          C.c_void_p(0x7000), patcher))
    h = C.c_void_p()
    return L.flsgpu_batch_plan_create(items, 1, 0, C.byref(h)), h


@pytest.mark.parametrize("field", ["right_packed", "left_packed", "exceptions_offsets", "counts",
                                   "exceptions", "positions"])
def test_batch_plan_rejects_null_alprd_arrays(field):
    """A batch item with a null array fails at plan creation with EINVAL, as
    the single-column entry point does, before anything touches the device."""
    ptrs = {f: 0x1000 * (i + 1) for i, f in enumerate(
        ["right_packed", "left_packed", "exceptions_offsets", "exceptions", "positions", "counts"])}
    ptrs[field] = None
    d = (C.c_uint16 * 8)(1, 2, 3, 0, 0, 0, 0, 0)
    col = A.ALPRDColumn(2048, 2, 50, 2, 3, d, ptrs["right_packed"], ptrs["left_packed"],
                        ptrs["exceptions_offsets"], ptrs["exceptions"], ptrs["positions"],
                        ptrs["counts"])
    for enc in (8, 9):  # ALP-RD f32 / f64
        col.right_bit_width = 18 if enc == 8 else 50
        rc, h = _plan(enc, col)
        assert rc == A.FLSGPU_EINVAL and not h.value, (enc, field)
        fn = A.lib().flsgpu_decompress_alprd_f32 if enc == 8 else A.lib().flsgpu_decompress_alprd_f64
        assert fn(C.byref(col), C.c_void_p(0x8000), 0, None) == A.FLSGPU_EINVAL


@pytest.mark.parametrize("field", ["exceptions", "positions", "counts", "offsets_counts",
                                   "bases", "factor_indices"])
def test_batch_plan_rejects_null_alp_arrays(field):
    bp = A.BPColumn(2048, 2, 0x1000, 0x2000, 0x3000)
    vals = dict(bases=0x4000, factor_indices=0x5000, fraction_indices=0x6000,
                exceptions_offsets=0x7000, exceptions=0x8000, positions=0x9000, counts=0xA000,
                offsets_counts=0xB000)
    vals[field] = None
    col = A.ALPColumn(A.FFORColumn(bp, vals["bases"]), vals["factor_indices"],
                      vals["fraction_indices"], 5, vals["exceptions_offsets"], vals["exceptions"],
                      vals["positions"], vals["counts"], vals["offsets_counts"])
    std_needs = field != "offsets_counts"
    ext_needs = field != "counts"
    # standard layout (enc 5, patcher stateful) and G-ALP (enc 7, patcher prefetch-all)
    for enc, patcher, needs in ((5, 3, std_needs), (7, 7, ext_needs)):
        rc, h = _plan(enc, col, patcher)
        if needs:
            assert rc == A.FLSGPU_EINVAL and not h.value, (enc, field)


def test_every_declared_entry_point_has_a_ctypes_signature():
    """Without argtypes ctypes passes a Python int as a 32-bit C int, which
    truncates 64-bit device pointers and sizes; every function the headers
    declare with parameters must have its signature set by some wrapper."""
    import re

    from flsgpu import alp as AL
    from flsgpu import batch as B
    from flsgpu import colfile as CF
    from flsgpu import encode_device as ED
    from flsgpu import selector as SEL
    L = A.lib()
    AL._enc_lib()
    B._sigs()
    CF._lib()
    ED._sigs()
    SEL._lib()
    no_params = set()
    import os
    inc = os.path.dirname(A.HEADER_PATH)
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            text = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, f)).read(), flags=re.S)
            no_params.update(re.findall(r"\b(flsgpu_[a-z0-9_]+)\s*\(\s*void\s*\)", text))
    missing = [n for n in A.header_symbols() if n not in no_params
               and getattr(getattr(L, n), "argtypes", None) is None]
    assert not missing, f"no ctypes signature: {missing}"
