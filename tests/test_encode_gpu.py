"""GPU parity of the device-side encoder: every array of the column built
in HBM equals the host encoder's (flsgpu_alp_encode_with_params_*, its G-ALP
conversion, FFOR packing at base = min), byte for byte, and decodes back to
the input bit for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import alp as AL  # noqa: E402
from flsgpu import device as D  # noqa: E402
from flsgpu import encode_device as ED  # noqa: E402
from flsgpu import selector as SEL  # noqa: E402

DEV = "cuda"
UI = {4: np.uint32, 8: np.uint64}


def bits(a):
    a = np.asarray(a)
    return a.view(UI[a.dtype.itemsize]) if a.dtype.kind == "f" else a


_SIGNED = {torch.uint16: (torch.int16, np.uint16), torch.uint32: (torch.int32, np.uint32),
           torch.uint64: (torch.int64, np.uint64)}


def t2n(t, n=None):
    if t.dtype in _SIGNED:
        st, nt = _SIGNED[t.dtype]
        a = t.view(st).cpu().numpy().view(nt)
    else:
        a = t.cpu().numpy()
    return a if n is None else a[:n]


def columns(dtype, seed):
    rng = np.random.default_rng(seed)
    nv = 24
    x = SEL.generate(nv, rng, 1024).astype(dtype).reshape(-1)
    v = x.reshape(nv, 1024)
    v[0] = np.nan                                    # every value an exception
    v[1, :1023] = np.inf                             # only slot 1023 encodes
    v[2, :5] = [np.nan, -0.0, np.inf, -np.inf, np.finfo(dtype).max]
    v[3] = np.finfo(dtype).tiny * np.arange(1024)    # subnormals and tiny values
    v[4, ::3] = np.float32(1e30) if dtype == np.float32 else 1e300
    v[5] = 0.0
    return x


def params(x, dtype, rng, mode):
    nv = len(x) // 1024
    if mode == "search":
        samples = np.ascontiguousarray(x.reshape(nv, 1024)[:, ::32])
        e, f, _ = AL.alp_search(samples)
    else:
        maxe = 10 if dtype == np.float32 else 18
        e = rng.integers(0, maxe + 1, nv).astype(np.uint8)
        f = np.minimum(rng.integers(0, maxe + 1, nv), e).astype(np.uint8)
        if dtype == np.float32:
            f = np.minimum(f, 9).astype(np.uint8)
    return e.astype(np.uint8), f.astype(np.uint8)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["search", "random"])
@pytest.mark.parametrize("galp", [False, True])
def test_alp_device_encode_matches_host(dtype, mode, galp):
    rng = np.random.default_rng(3)
    x = columns(dtype, 11)
    if galp:  # vectors with a lane of 64 exceptions cannot be expressed in G-ALP
        v = x.reshape(-1, 1024)
        v[0] = np.round(v[6], 2)
        v[1] = np.round(v[7], 1)
        if dtype == np.float64:
            v[3] = np.round(v[8], 2)
    e, f = params(x, dtype, rng, mode)
    host = AL.alp_encode_with_params(x, e, f)
    xd = torch.from_numpy(x).to(DEV)
    args = (xd, torch.from_numpy(e).to(DEV), torch.from_numpy(f).to(DEV))
    if galp:
        try:
            host = AL.galp_convert(host)
        except A.FlsgpuError:
            # random f64 parameters leave whole lanes (64 slots) as exceptions,
            # which the 16-bit G-ALP lane header cannot express: the device
            # encoder must refuse the column as the host converter does
            assert mode == "random"
            with pytest.raises(A.FlsgpuError):
                ED.alp_encode_device(*args, galp=True)
            return
    col, info = ED.alp_encode_device(*args, galp=galp)
    nw, ne = info["packed_words"], info["exceptions"]
    assert nw == len(host.packed) and ne == len(host.exceptions)
    assert np.array_equal(t2n(col.packed, nw), host.packed)
    for name in ("bit_widths", "vector_offsets", "bases", "factor_indices", "fraction_indices",
                 "exceptions_offsets"):
        assert np.array_equal(t2n(getattr(col, name)), getattr(host, name)), name
    assert np.array_equal(bits(t2n(col.exceptions, ne)), bits(host.exceptions))
    assert np.array_equal(t2n(col.positions, ne), host.positions)
    assert np.array_equal(t2n(info["counts"]), host.counts)
    if galp:
        assert np.array_equal(t2n(col.offsets_counts), host.offsets_counts)
    out = AL.decompress(col).cpu().numpy()
    assert np.array_equal(bits(out), bits(x))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_out_of_range_parameters_encode_as_zero_zero(dtype):
    """(exponent, factor) pairs outside the tables (exponent above 10 / 18,
    factor above the exponent, float factor 10) encode their vector with
    (0, 0): the same bytes as the host encoder at (0, 0), lossless."""
    rng = np.random.default_rng(8)
    x = np.round(rng.uniform(-100, 100, 4 * 1024), 1).astype(dtype)
    x[::7] = np.round(x[::7])  # some integers, which (0, 0) encodes without exceptions
    maxe = 10 if dtype == np.float32 else 18
    e = np.array([maxe + 1, 3, 255, 10], np.uint8)
    f = np.array([0, 4, 0, 10], np.uint8)   # factor > exponent; factor 10 (floats)
    ok = np.array([False, False, False, dtype == np.float64])
    e0 = np.where(ok, e, 0).astype(np.uint8)
    f0 = np.where(ok, f, 0).astype(np.uint8)
    host = AL.alp_encode_with_params(x, e0, f0)
    xd = torch.from_numpy(x).to(DEV)
    col, info = ED.alp_encode_device(xd, torch.from_numpy(e).to(DEV), torch.from_numpy(f).to(DEV))
    assert np.array_equal(t2n(col.factor_indices), f0)
    assert np.array_equal(t2n(col.fraction_indices), e0)
    nw, ne = info["packed_words"], info["exceptions"]
    assert nw == len(host.packed) and ne == len(host.exceptions)
    assert np.array_equal(t2n(col.packed, nw), host.packed)
    assert np.array_equal(bits(t2n(col.exceptions, ne)), bits(host.exceptions))
    assert np.array_equal(bits(AL.decompress(col).cpu().numpy()), bits(x))


def test_galp_lane_overflow_is_reported():
    x = np.round(np.random.default_rng(1).uniform(-100, 100, 2048), 2)
    x[1024 + 3::16] = np.pi          # 64 exceptions in lane 3 of vector 1
    e = np.full(2, 2, np.uint8)
    f = np.zeros(2, np.uint8)
    xd = torch.from_numpy(x).to(DEV)
    args = (xd, torch.from_numpy(e).to(DEV), torch.from_numpy(f).to(DEV))
    col, info = ED.alp_encode_device(*args, galp=False)       # fine in the standard layout
    assert info["exceptions"] >= 64
    assert np.array_equal(bits(AL.decompress(col).cpu().numpy()), bits(x))
    with pytest.raises(A.FlsgpuError):
        ED.alp_encode_device(*args, galp=True)


@pytest.mark.parametrize("width", [32, 64])
def test_ffor_device_encode_matches_host(width):
    rng = np.random.default_rng(width)
    ut = UI[width // 8]
    nv = 40
    bws = rng.integers(0, width + 1, nv)
    bws[:3] = [0, width, 1]
    lo = np.frombuffer(rng.bytes(nv * width // 8), dtype=ut).copy()
    span = np.array([(1 << int(b)) - 1 if b < width else (1 << width) - 1 for b in bws], dtype=object)
    vals = np.empty((nv, 1024), ut)
    for v in range(nv):
        off = np.frombuffer(rng.bytes(1024 * width // 8), dtype=ut) & ut(int(span[v]))
        base = ut(0) if bws[v] == width else ut(min(int(lo[v]), (1 << width) - 1 - int(span[v])))
        vals[v] = off + base
    vals = vals.reshape(-1)
    mn = vals.reshape(nv, 1024).min(axis=1)
    mx = vals.reshape(nv, 1024).max(axis=1)
    exp_bw = np.array([int(int(a) - int(b)).bit_length() for a, b in zip(mx, mn)], np.uint8)
    packed, offs = AL.ffor_pack(vals, exp_bw, mn)
    col, words = ED.ffor_encode_device(torch.from_numpy(vals).to(DEV))
    assert words == len(packed)
    assert np.array_equal(t2n(col.bit_widths), exp_bw)
    assert np.array_equal(t2n(col.bases), mn)
    assert np.array_equal(t2n(col.vector_offsets), offs)
    assert np.array_equal(t2n(col.packed, words), packed)
    assert np.array_equal(t2n(D.decompress(col)), vals)
