"""GPU parity of the ALP / G-ALP kernels against the CPU oracle.

Mirrors the reference's ALP differential test (src/test.cu:197-242,
test-alp.sh): random FFOR data with bases in [2, 20], factor / fraction
indices in the lower halves of their tables, random exception bytes at
random sorted positions (src/engine/data.cuh:436-505), every patcher of both
layouts, bitwise comparison. Encoded real-shaped data must also round-trip.
"""
import numpy as np
import pytest

import oracle_py as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import alp as AL  # noqa: E402

STD_PATCHERS = ["stateless", "stateful"]
GALP_PATCHERS = ["naive", "naive-branchless", "prefetch-position", "prefetch-all",
                 "prefetch-all-branchless"]
FT = {32: np.float32, 64: np.float64}
UT = {32: np.uint32, 64: np.uint64}


def rand_bits(rng, n, width):
    return np.frombuffer(rng.bytes(n * width // 8), dtype=UT[width]).copy()


def reference_style_column(rng, width, n_vecs, bws, max_exc):
    """generate_alp_column + modify_alp_exception_count (data.cuh:436-505)."""
    nl = 1024 // width
    bws = np.asarray(bws, np.uint8)
    offs = np.zeros(n_vecs, np.uint64)
    offs[1:] = np.cumsum(bws.astype(np.uint64) * np.uint64(nl))[:-1]
    packed = rand_bits(rng, int(bws.astype(np.int64).sum()) * nl, width)
    bases = rng.integers(2, 21, n_vecs).astype(UT[width])
    n_fact, n_frac = (10, 11) if width == 32 else (19, 21)
    fi = rng.integers(0, n_fact // 2 + 1, n_vecs).astype(np.uint8)
    ri = rng.integers(0, n_frac // 2 + 1, n_vecs).astype(np.uint8)
    counts = rng.integers(0, max_exc + 1, n_vecs).astype(np.uint16)
    eo = np.zeros(n_vecs, np.uint64)
    eo[1:] = np.cumsum(counts.astype(np.uint64))[:-1]
    exc = rand_bits(rng, int(counts.sum()), width).view(FT[width])
    pos = np.concatenate([np.sort(rng.choice(1024, int(c), replace=False)) for c in counts]
                         ).astype(np.uint16) if counts.sum() else np.zeros(0, np.uint16)
    return AL.AlpHostColumn(n_vecs * 1024, packed, bws, offs, bases, fi, ri, eo, exc, pos, counts)


def oracle_decode(col, patch=True):
    ocol = O.ALPColumn(col.n_values, O.FFORColumn(col.n_values, col.packed if len(col.packed) else
                                                  np.zeros(1, col.packed.dtype), col.bit_widths,
                                                  col.vector_offsets, col.bases),
                       col.factor_indices, col.fraction_indices, col.exceptions_offsets,
                       col.exceptions if len(col.exceptions) else np.zeros(1, col.dtype),
                       col.positions if len(col.positions) else np.zeros(1, np.uint16),
                       col.counts if patch else np.zeros_like(col.counts))
    return O.alp_column_decode(ocol)


def bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("vname,variant", sorted(A.VARIANTS.items()))
def test_reference_style_every_width(width, vname, variant):
    rng = np.random.default_rng(500 + width + variant)
    for bw in range(width + 1):
        col = reference_style_column(rng, width, 6, [bw] * 6, 10)
        exp = oracle_decode(col)
        d = AL.DeviceAlpColumn.from_host(col)
        for p in STD_PATCHERS:
            got = AL.decompress(d, patcher=A.PATCHERS[p], variant=variant).cpu().numpy()
            assert np.array_equal(bits(got), bits(exp)), f"{vname} f{width} bw={bw} {p}"
        g = AL.DeviceAlpColumn.from_host(AL.galp_convert(col))
        for p in GALP_PATCHERS:
            got = AL.decompress(g, patcher=A.PATCHERS[p], variant=variant).cpu().numpy()
            assert np.array_equal(bits(got), bits(exp)), f"{vname} f{width} bw={bw} G-ALP {p}"


@pytest.mark.parametrize("width", [32, 64])
def test_dummy_patcher_skips_exceptions(width):
    rng = np.random.default_rng(77)
    col = reference_style_column(rng, width, 9, rng.integers(0, 17, 9), 10)
    d = AL.DeviceAlpColumn.from_host(col)
    for p in ("dummy", "none"):
        got = AL.decompress(d, patcher=A.PATCHERS[p]).cpu().numpy()
        assert np.array_equal(bits(got), bits(oracle_decode(col, patch=False)))


@pytest.mark.parametrize("width", [32, 64])
def test_many_exceptions_per_lane(width):
    """Dense exceptions (up to 63 in one lane) for both layouts."""
    rng = np.random.default_rng(78)
    col = reference_style_column(rng, width, 4, [7, 0, 13, width], 0)
    nl = 1024 // width
    pos_l, cnt = [], []
    for v in range(4):
        lane = v % nl
        slots = np.arange(min(63, width))
        p = np.sort(lane + nl * slots).astype(np.uint16)
        others = np.arange(1024)[np.arange(1024) % nl != lane]
        extra = np.sort(rng.choice(others, 100, replace=False))
        p = np.sort(np.concatenate([p, extra])).astype(np.uint16)
        pos_l.append(p)
        cnt.append(len(p))
    col.counts = np.array(cnt, np.uint16)
    col.exceptions_offsets = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint64)
    col.positions = np.concatenate(pos_l)
    col.exceptions = rand_bits(rng, len(col.positions), width).view(FT[width])
    exp = oracle_decode(col)
    d = AL.DeviceAlpColumn.from_host(col)
    g = AL.DeviceAlpColumn.from_host(AL.galp_convert(col))
    for variant in (0, A.VARIANT_HALF_LANE_STAGED, A.VARIANT_QUARTER_LANE_STAGED,
                    A.VARIANT_EIGHTH_LANE_STAGED):
        got = AL.decompress(d, variant=variant).cpu().numpy()
        assert np.array_equal(bits(got), bits(exp)), f"variant {variant}"
        got = AL.decompress(g, variant=variant).cpu().numpy()
        assert np.array_equal(bits(got), bits(exp)), f"G-ALP variant {variant}"


@pytest.mark.parametrize("width", [32, 64])
def test_staged_variants_misaligned_output(width):
    """The LDS-staged variants need a 16-byte aligned output; any other
    output pointer takes the per-thread store path with the same result."""
    rng = np.random.default_rng(79)
    col = reference_style_column(rng, width, 7, rng.integers(0, width + 1, 7), 10)
    exp = oracle_decode(col)
    d = AL.DeviceAlpColumn.from_host(col)
    dt = torch.float32 if width == 32 else torch.float64
    for variant in (A.VARIANT_HALF_LANE_STAGED, A.VARIANT_QUARTER_LANE_STAGED,
                    A.VARIANT_EIGHTH_LANE_STAGED):
        buf = torch.full((col.n_vecs * 1024 + 1,), -1.0, dtype=dt, device="cuda")
        out = buf[1:]
        assert out.data_ptr() % 16 != 0
        AL.decompress(d, out=out, variant=variant)
        assert np.array_equal(bits(out.cpu().numpy()), bits(exp))
        assert buf[0].item() == -1.0


def c3_like(width, n_vecs, rng):
    if width == 64:
        x = rng.integers(-1_000_000, 1_000_001, n_vecs * 1024) / 100.0
        repl = rng.uniform(-1e4, 1e4, (n_vecs, 10))
    else:
        x = (rng.integers(-10_000, 10_001, n_vecs * 1024).astype(np.float32) / np.float32(10.0))
        repl = rng.uniform(-1e3, 1e3, (n_vecs, 10)).astype(np.float32)
    v = x.reshape(n_vecs, 1024)
    for i in range(n_vecs):
        v[i, rng.choice(1024, 10, replace=False)] = repl[i]
    return x


@pytest.mark.parametrize("width", [32, 64])
def test_encoded_roundtrip_and_specials(width):
    rng = np.random.default_rng(90 + width)
    x = c3_like(width, 40, rng)
    x[5] = np.nan
    x[77] = np.inf
    x[1030] = -np.inf
    x[2000] = -0.0
    x[3000] = np.finfo(FT[width]).max
    x[4000] = np.finfo(FT[width]).tiny / 4  # subnormal
    col = AL.alp_encode(x)
    assert col is not None
    for c in (col, AL.galp_convert(col)):
        got = AL.decompress(AL.DeviceAlpColumn.from_host(c)).cpu().numpy()
        assert np.array_equal(bits(got), bits(x))


@pytest.mark.parametrize("width", [32, 64])
def test_query_any(width):
    rng = np.random.default_rng(91)
    x = c3_like(width, 16, rng)
    col = AL.alp_encode(x)
    for c in (col, AL.galp_convert(col)):
        d = AL.DeviceAlpColumn.from_host(c)
        exc_val = c.exceptions[3]
        assert AL.query_any(d, float(exc_val)).item() == 1          # value only present as exception
        assert AL.query_any(d, float(x[123])).item() == 1
        assert AL.query_any(d, 0.3214).item() == int((x == FT[width](0.3214)).any())
        assert AL.query_any(d, 1e30).item() == 0
    # exceptions replace the decoded value: vector 0 decodes to 3.0 everywhere
    # but every slot is an exception (7.5); vector 1 decodes to 5.0
    col = reference_style_column(rng, width, 2, [0, 0], 0)
    col.bases[:] = [3, 5]
    col.factor_indices[:] = 0
    col.fraction_indices[:] = 0
    col.counts = np.array([1024, 0], np.uint16)
    col.exceptions_offsets = np.array([0, 1024], np.uint64)
    col.positions = np.arange(1024, dtype=np.uint16)
    col.exceptions = np.full(1024, 7.5, FT[width])
    d = AL.DeviceAlpColumn.from_host(col)
    assert AL.query_any(d, 3.0, patcher=A.PATCHERS["stateless"]).item() == 0
    assert AL.query_any(d, 3.0, patcher=A.PATCHERS["dummy"]).item() == 1
    assert AL.query_any(d, 7.5, patcher=A.PATCHERS["stateful"]).item() == 1
    assert AL.query_any(d, 5.0).item() == 1


@pytest.mark.parametrize("width", [32, 64])
def test_query_any_against_oracle(width):
    """query_any at every kind of needle, against the oracle's decode: bases
    small, negative and close to the signed maximum (where (base + v) * FACT
    can wrap, so the kernel decodes every value instead of searching the
    interval of matching offsets), widths 0 to W; needles taken from decoded
    values (the interval's ends included), from exceptions and from between
    two decoded values, plus NaN, infinities and signed zeros."""
    rng = np.random.default_rng(93)
    smax = (1 << (width - 1)) - 1
    for bw in (0, 1, 2, 5, 13, 21, width - 2, width - 1, width):
        for kind in ("small", "negative", "near_max"):
            n_vecs = 4
            col = reference_style_column(rng, width, n_vecs, [bw] * n_vecs, 6)
            if kind == "negative":
                col.bases[:] = np.array([-1000, -3, -(1 << 20), -7], np.int64).astype(UT[width])
            elif kind == "near_max":
                col.bases[:] = np.array([smax - 5, smax // 3, smax // 10 ** 3, 7], dtype=UT[width])
            raw = oracle_decode(col, patch=False)
            exp_vals = oracle_decode(col)
            needles = [raw[int(i)] for i in rng.choice(len(raw), 6, replace=False)]
            for v in range(n_vecs):  # each vector's smallest and largest decoded value
                blk = raw[v * 1024:(v + 1) * 1024]
                fin = blk[np.isfinite(blk)]
                if len(fin):
                    needles += [fin.min(), fin.max()]
            if len(col.exceptions):
                needles.append(col.exceptions[0])
            srt = np.unique(raw[np.isfinite(raw)])
            if len(srt) > 1:
                needles.append(FT[width]((float(srt[0]) + float(srt[1])) / 2))
            needles += [FT[width](np.nan), FT[width](np.inf), FT[width](-np.inf), FT[width](0.0),
                        FT[width](-0.0)]
            for layout_col in (col, AL.galp_convert(col)):
                d = AL.DeviceAlpColumn.from_host(layout_col)
                for nd in needles:
                    want = int(bool((exp_vals == nd).any()))
                    got = AL.query_any(d, float(nd)).item()
                    assert got == want, (width, bw, kind, float(nd))


@pytest.mark.parametrize("width", [32, 64])
def test_multi_column(width):
    rng = np.random.default_rng(92)
    n_vecs = 4
    nl = 1024 // width
    cols = []
    for c in range(10):
        x = (rng.integers(-4000, 4000, n_vecs * 1024) / 4).astype(FT[width])
        x[rng.choice(n_vecs * 1024, 20, replace=False)] = rng.uniform(-1, 1, 20).astype(FT[width])
        x[2 * 1024 + 3 + nl * np.arange(width)] = FT[width](12.5)
        cols.append(AL.alp_encode(x))
    for layout in ("std", "galp"):
        dc = [AL.DeviceAlpColumn.from_host(c if layout == "std" else AL.galp_convert(c)) for c in cols]
        for n in (1, 2, 5, 10):
            assert AL.query_multi_column(dc[:n], 12.5).item() == 1
            assert AL.query_multi_column(dc[:n], 13.5).item() == 0
