"""GPU parity of the device-side decoder header (include/flsgpu_device.hpp).

Test kernels in tests/cpp/device_api.hip follow the reference's
decompress_column / query_column shapes (src/engine/kernels.cuh:13-54) with
BPDecompressor, FFORDecompressor and ALPDecompressor (fls.cuh:620-650,
alp.cuh:463-485) for UNPACK_N_VECTORS in {1, 4} and UNPACK_N_VALUES in
{1, 2, 4}; every output is compared bit for bit with the CPU oracle.
"""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import alp as AL  # noqa: E402

import test_alp_gpu as TA  # noqa: E402
import test_ffor_gpu as TF  # noqa: E402

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build",
                        "libdevice_api.so")
SHAPES = {0: "1x1", 1: "1x4", 2: "4x1", 3: "4x2"}  # vectors x values per call
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C tests/cpp)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        for w in (32, 64):
            f = getattr(L, f"dapi_decompress_u{w}")
            f.argtypes = [C.POINTER(A.FFORColumn), P, C.c_int, P]
            f.restype = C.c_int
            f = getattr(L, f"dapi_decompress_alp_f{w}")
            f.argtypes = [C.POINTER(A.ALPColumn), P, C.c_int, C.c_int, P]
            f.restype = C.c_int
        L.dapi_query_any_ffor_u32.argtypes = [C.POINTER(A.FFORColumn), C.c_uint32, P, P]
        L.dapi_query_any_ffor_u32.restype = C.c_int
        L.dapi_query_any_alp_f64.argtypes = [C.POINTER(A.ALPColumn), C.c_double, P, P]
        L.dapi_query_any_alp_f64.restype = C.c_int
        _lib = L
    return _lib


def decode_ffor(col, width, shape):
    out = torch.full((col.n_values,), 0xAB, dtype=torch.uint32 if width == 32 else torch.uint64,
                     device="cuda")
    fn = getattr(lib(), f"dapi_decompress_u{width}")
    assert fn(C.byref(col.ffor_struct()), out.data_ptr(), shape, None) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


def decode_alp(d, width, layout, shape):
    dt = torch.float32 if width == 32 else torch.float64
    out = torch.full((d.n_values,), 7.0, dtype=dt, device="cuda")
    fn = getattr(lib(), f"dapi_decompress_alp_f{width}")
    assert fn(C.byref(d.struct()), out.data_ptr(), layout, shape, None) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_bp_ffor_every_width(width, shape):
    rng = np.random.default_rng(300 + width + shape)
    n_vecs = 2 * (width + 1) + 3  # every width twice, ragged against 4-vector groups
    widths = np.concatenate([np.arange(width + 1), rng.permutation(width + 1), [0, width, 1]])
    for with_bases in (True, False):
        packed, bws, offs, bases = TF.random_column(rng, n_vecs, width, widths, with_bases)
        col = TF.gpu_column(packed, bws, offs, bases, width)
        got = decode_ffor(col, width, shape)
        exp = TF.oracle_decode(packed, bws, offs, bases, width)
        assert np.array_equal(got, exp), f"u{width} {SHAPES[shape]} bases={with_bases}"


@pytest.mark.parametrize("width", [32, 64])
def test_invalid_width_decodes_zeros(width):
    rng = np.random.default_rng(5)
    packed, bws, offs, bases = TF.random_column(rng, 5, width, [9] * 5)
    bad = bws.copy()
    bad[2] = width + 1
    col = TF.gpu_column(packed, bad, offs, bases, width)
    exp = TF.oracle_decode(packed, bws, offs, bases, width)
    for shape in SHAPES:
        got = decode_ffor(col, width, shape)
        assert not got[2048:3072].any()
        assert np.array_equal(np.delete(got, np.s_[2048:3072]), np.delete(exp, np.s_[2048:3072]))


def shuffle_exceptions(col, rng):
    """Same exceptions, positions in random order within each vector."""
    pos, exc = col.positions.copy(), col.exceptions.copy()
    for v in range(col.n_vecs):
        a = int(col.exceptions_offsets[v])
        b = a + int(col.counts[v])
        p = rng.permutation(b - a)
        pos[a:b], exc[a:b] = pos[a:b][p], exc[a:b][p]
    col.positions, col.exceptions = pos, exc
    return col


@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_alp_layouts(width, shape):
    rng = np.random.default_rng(400 + width + shape)
    n_vecs = 11
    col = TA.reference_style_column(rng, width, n_vecs, rng.integers(0, width + 1, n_vecs), 40)
    if shape % 2:
        col = shuffle_exceptions(col, rng)
    exp = TA.oracle_decode(col)
    d = AL.DeviceAlpColumn.from_host(col)
    g = AL.DeviceAlpColumn.from_host(AL.galp_convert(col))
    assert np.array_equal(TA.bits(decode_alp(d, width, 0, shape)), TA.bits(exp)), "standard"
    assert np.array_equal(TA.bits(decode_alp(g, width, 1, shape)), TA.bits(exp)), "G-ALP"
    unpatched = TA.oracle_decode(col, patch=False)
    assert np.array_equal(TA.bits(decode_alp(d, width, 2, shape)), TA.bits(unpatched)), "dummy"


@pytest.mark.parametrize("width", [32, 64])
def test_alp_encoded_column_matches_library(width):
    """A column from the ALP encoder: the header's decode equals the
    library's decode and the original values."""
    rng = np.random.default_rng(41)
    x = TA.c3_like(width, 24, rng)
    col = AL.alp_encode(x)
    d = AL.DeviceAlpColumn.from_host(col)
    lib_out = AL.decompress(d).cpu().numpy()
    for shape in SHAPES:
        got = decode_alp(d, width, 0, shape)
        assert np.array_equal(TA.bits(got), TA.bits(lib_out))
        assert np.array_equal(TA.bits(got), TA.bits(x))


def test_query_shapes():
    rng = np.random.default_rng(8)
    packed, bws, offs, bases = TF.random_column(rng, 9, 32, [13] * 9)
    col = TF.gpu_column(packed, bws, offs, bases, 32)
    dec = TF.oracle_decode(packed, bws, offs, bases, 32)
    L = lib()
    for magic, want in ((int(dec[4321]), True), (int(dec.max()) + 1, False)):
        found = torch.zeros(1, dtype=torch.bool, device="cuda")
        assert L.dapi_query_any_ffor_u32(C.byref(col.ffor_struct()), magic, found.data_ptr(),
                                         None) == 0
        assert bool(found.item()) == want
    x = TA.c3_like(64, 6, rng)
    g = AL.DeviceAlpColumn.from_host(AL.galp_convert(AL.alp_encode(x)))
    for magic, want in ((float(x[1030]), True), (123456.789, bool(np.any(x == 123456.789)))):
        found = torch.zeros(1, dtype=torch.bool, device="cuda")
        assert L.dapi_query_any_alp_f64(C.byref(g.struct()), magic, found.data_ptr(), None) == 0
        assert bool(found.item()) == want
