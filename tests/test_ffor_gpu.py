"""GPU parity of the FastLanes / FFOR kernels against the CPU oracle.

Mirrors the reference's differential test (src/test.cu:84-117, test-fls.sh):
random packed data and bases (data.cuh:379-389), every bit width 0..W, every
unpacker variant, bitwise comparison (verification.cuh:52-75). Queries and
sums are checked against closed forms on the oracle's decoded output.
"""
import numpy as np
import pytest

import oracle_py as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import device as D  # noqa: E402

VARIANTS = sorted(A.VARIANTS.items())
UT = {32: np.uint32, 64: np.uint64}


def rand_bits(rng, n, width):
    """n uniformly random W-bit words (every bit pattern possible)."""
    return np.frombuffer(rng.bytes(n * width // 8), dtype=UT[width]).copy()


def random_column(rng, n_vecs, width, widths, with_bases=True):
    """Random packed words with per-vector bit widths (the reference's
    generate_random_ffor_column, src/engine/data.cuh:333-389)."""
    nl = 1024 // width
    bws = np.asarray(widths, dtype=np.uint8)
    offs = np.zeros(n_vecs, dtype=np.uint64)
    if n_vecs:
        offs[1:] = np.cumsum(bws.astype(np.uint64) * np.uint64(nl))[:-1]
    n_words = int(bws.astype(np.int64).sum()) * nl
    packed = rand_bits(rng, n_words, width)
    bases = rand_bits(rng, n_vecs, width) if with_bases else None
    return packed, bws, offs, bases


def oracle_decode(packed, bws, offs, bases, width):
    col = O.FFORColumn(len(bws) * 1024, packed if len(packed) else np.zeros(1, UT[width]), bws, offs,
                       bases if bases is not None else np.zeros(len(bws), UT[width]))
    return O.ffor_column_decode(col)


def gpu_column(packed, bws, offs, bases, width):
    return D.DeviceFFORColumn.from_host(packed, bws, offs, bases, len(bws) * 1024)


@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("vname,variant", VARIANTS)
def test_decompress_every_width_uniform(width, vname, variant):
    rng = np.random.default_rng(1000 + width + variant)
    for bw in range(width + 1):
        n_vecs = 5  # not a multiple of the vectors per wavefront: ragged tail
        packed, bws, offs, bases = random_column(rng, n_vecs, width, [bw] * n_vecs)
        col = gpu_column(packed, bws, offs, bases, width)
        got = D.decompress(col, variant=variant).cpu().numpy()
        exp = oracle_decode(packed, bws, offs, bases, width)
        assert np.array_equal(got, exp), f"{vname} u{width} bw={bw}"


@pytest.mark.parametrize("width", [32, 64])
@pytest.mark.parametrize("vname,variant", VARIANTS)
def test_decompress_mixed_widths_and_bp(width, vname, variant):
    rng = np.random.default_rng(7 * width + variant)
    n_vecs = 3 * (width + 1) + 1
    widths = rng.integers(0, width + 1, n_vecs)
    for with_bases in (True, False):
        packed, bws, offs, bases = random_column(rng, n_vecs, width, widths, with_bases)
        col = gpu_column(packed, bws, offs, bases, width)
        got = D.decompress(col, variant=variant).cpu().numpy()
        exp = oracle_decode(packed, bws, offs, bases, width)
        assert np.array_equal(got, exp), f"{vname} u{width} bases={with_bases}"


@pytest.mark.parametrize("width", [32, 64])
def test_filter_sum_mixed_widths(width):
    """Widths that differ between vectors (the fused kernel's per-unit path
    next to its batched same-width path) and a vector with a width above W,
    which contributes no rows."""
    rng = np.random.default_rng(14)
    m64 = (1 << 64) - 1
    n_vecs = 67
    widths = np.full(n_vecs, 11, np.int64)
    widths[20:45] = rng.integers(0, width + 1, 25)  # mixed block in the middle
    packed, bws, offs, bases = random_column(rng, n_vecs, width, widths)
    bases = rng.integers(0, 1 << 20, n_vecs).astype(UT[width])
    bad = bws.copy()
    bad[30] = width + 1
    col = gpu_column(packed, bad, offs, bases, width)
    dec = oracle_decode(packed, bws, offs, bases, width).reshape(n_vecs, 1024)
    keep = np.ones(n_vecs, bool)
    keep[30] = False
    off = dec - bases[:, None]
    half = np.array([(1 << (int(b) - 1)) if b > 0 else 0 for b in bws], dtype=np.uint64)
    sel = (off.astype(np.uint64) < half[:, None]) & keep[:, None]
    r = D.query_filter_sum(col, A.PRED_OFFSET_LOWER_HALF).cpu().numpy()
    assert int(r[0]) == int(sel.sum())
    assert int(r[1]) == sum(int(x) for x in dec[sel]) & m64
    r = D.query_filter_sum(col, A.PRED_ALL).cpu().numpy()
    assert int(r[0]) == 1024 * int(keep.sum())
    assert int(r[1]) == sum(int(x) for x in dec[keep].reshape(-1)) & m64


@pytest.mark.parametrize("width", [32, 64])
def test_staged_variants_misaligned_output(width):
    """A staged variant with an output that is not 16-byte aligned takes the
    per-thread store path; the words around the output stay untouched."""
    rng = np.random.default_rng(21 + width)
    n_vecs = 9
    packed, bws, offs, bases = random_column(rng, n_vecs, width, rng.integers(0, width + 1, n_vecs))
    col = gpu_column(packed, bws, offs, bases, width)
    exp = oracle_decode(packed, bws, offs, bases, width)
    dt = torch.uint32 if width == 32 else torch.uint64
    for variant in (A.VARIANT_HALF_LANE_STAGED, A.VARIANT_QUARTER_LANE_STAGED,
                    A.VARIANT_EIGHTH_LANE_STAGED):
        buf = torch.zeros(n_vecs * 1024 + 2, dtype=dt, device="cuda")
        out = buf[1:-1]
        assert out.data_ptr() % 16 != 0
        D.decompress(col, out=out, variant=variant)
        assert np.array_equal(out.cpu().numpy(), exp)
        assert int(buf[0].item()) == 0 and int(buf[-1].item()) == 0


@pytest.mark.parametrize("width", [32, 64])
def test_decompress_edges(width):
    rng = np.random.default_rng(3)
    # empty column
    packed, bws, offs, bases = random_column(rng, 0, width, [])
    col = gpu_column(packed, bws, offs, bases, width)
    assert D.decompress(col).numel() == 0
    # single vector (partial wavefront), full width and zero width
    for bw in (0, width):
        packed, bws, offs, bases = random_column(rng, 1, width, [bw])
        col = gpu_column(packed, bws, offs, bases, width)
        got = D.decompress(col).cpu().numpy()
        assert np.array_equal(got, oracle_decode(packed, bws, offs, bases, width))
    # bit width above W decodes as zeros, neighbours unaffected
    packed, bws, offs, bases = random_column(rng, 3, width, [5, 5, 5])
    bws_bad = bws.copy()
    bws_bad[1] = width + 1
    col = gpu_column(packed, bws_bad, offs, bases, width)
    got = D.decompress(col).cpu().numpy()
    exp = oracle_decode(packed, bws, offs, bases, width)
    assert np.array_equal(got[:1024], exp[:1024]) and np.array_equal(got[2048:], exp[2048:])
    assert not got[1024:2048].any()


@pytest.mark.parametrize("width", [32, 64])
def test_query_any(width):
    rng = np.random.default_rng(11)
    for bw in (0, 1, 7, 13, width // 2, width - 1, width):
        packed, bws, offs, bases = random_column(rng, 9, width, [bw] * 9)
        col = gpu_column(packed, bws, offs, bases, width)
        dec = oracle_decode(packed, bws, offs, bases, width)
        present = dec[rng.integers(0, len(dec))]
        absent = UT[width](int(dec.max()) + 1) if int(dec.max()) < 2 ** width - 1 else None
        assert D.query_any(col, int(present)).item() == 1
        if absent is not None and not (dec == absent).any():
            assert D.query_any(col, int(absent)).item() == 0
    # reference FFOR query data: all-zero column, magic 1 absent (data.cuh:393-426)
    zeros = np.zeros(4 * 1024, dtype=UT[width])
    from flsgpu import encode as E
    p, b, o, bs = E.ffor_encode_uniform(zeros, 3)
    col = gpu_column(p, b, o, bs, width)
    assert D.query_any(col, 1).item() == 0
    assert D.query_any(col, 0).item() == 1


@pytest.mark.parametrize("width", [32, 64])
def test_filter_sum(width):
    """Count and sum of the rows passing each predicate, against closed forms
    on the oracle's decode; bases both small (the algebraic no-wrap path) and
    full-range (offset + base wraps: the per-value path)."""
    rng = np.random.default_rng(12)
    m64 = (1 << 64) - 1
    for bw in (0, 1, 2, 13, 27, 28, width - 1, width):
        for small_bases in (True, False):
            n_vecs = 37
            packed, bws, offs, bases = random_column(rng, n_vecs, width, [bw] * n_vecs)
            if small_bases:
                bases = rng.integers(0, 1 << 20, n_vecs).astype(UT[width])
            col = gpu_column(packed, bws, offs, bases, width)
            dec = oracle_decode(packed, bws, offs, bases, width)
            off = (dec.reshape(n_vecs, 1024) - bases[:, None]).reshape(-1)
            tag = f"u{width} bw={bw} small_bases={small_bases}"
            # offset predicate: offset < 2^(bw-1), none pass at bw 0
            sel = (off < (1 << (bw - 1))) if bw > 0 else np.zeros(len(dec), bool)
            r = D.query_filter_sum(col, A.PRED_OFFSET_LOWER_HALF).cpu().numpy()
            assert int(r[0]) == int(sel.sum()), tag
            assert int(r[1]) == sum(int(x) for x in dec[sel]) & m64, tag
            # every row
            r = D.query_filter_sum(col, A.PRED_ALL).cpu().numpy()
            assert int(r[0]) == len(dec) and int(r[1]) == sum(int(x) for x in dec) & m64, tag
            # value range predicate, including empty and full ranges
            for lo, hi in (sorted(int(x) for x in rng.choice(dec, 2)), (0, 0),
                           (0, (1 << width) - 1), (int(dec.max()), int(dec.min()))):
                sel = (dec >= UT[width](lo)) & (dec < UT[width](hi))
                r = D.query_filter_sum(col, A.PRED_RANGE, lo, hi).cpu().numpy()
                assert int(r[0]) == int(sel.sum()), (tag, lo, hi)
                assert int(r[1]) == sum(int(x) for x in dec[sel]) & m64, (tag, lo, hi)


def _sum_mod64(x):
    """Sum mod 2^64 of unsigned values (numpy's uint64 sum wraps)."""
    return int(np.asarray(x, dtype=np.uint64).sum(dtype=np.uint64))


@pytest.mark.parametrize("width", [32, 64])
def test_filter_sum_every_width(width):
    """Every bit width 0..W through the fused kernel's batched path: columns of
    64 vectors (every wave's group of units full) and of 37 (a ragged last
    group), bases small (offset + base never wraps) and full-range (wrapping
    32-bit values take the per-value correction), each predicate against
    closed forms on the oracle's decode."""
    rng = np.random.default_rng(21)
    for bw in range(width + 1):
        for n_vecs, small_bases in ((64, True), (37, False), (64, False)):
            packed, bws, offs, bases = random_column(rng, n_vecs, width, [bw] * n_vecs)
            if small_bases:
                bases = rng.integers(0, 1 << 20, n_vecs).astype(UT[width])
            col = gpu_column(packed, bws, offs, bases, width)
            dec = oracle_decode(packed, bws, offs, bases, width)
            off = (dec.reshape(n_vecs, 1024) - bases[:, None]).reshape(-1)
            tag = f"u{width} bw={bw} n_vecs={n_vecs} small_bases={small_bases}"
            sel = (off < (1 << (bw - 1))) if bw > 0 else np.zeros(len(dec), bool)
            r = D.query_filter_sum(col, A.PRED_OFFSET_LOWER_HALF).cpu().numpy()
            assert (int(r[0]), int(r[1])) == (int(sel.sum()), _sum_mod64(dec[sel])), tag
            r = D.query_filter_sum(col, A.PRED_ALL).cpu().numpy()
            assert (int(r[0]), int(r[1])) == (len(dec), _sum_mod64(dec)), tag
            lo, hi = sorted(int(x) for x in rng.choice(dec, 2))
            sel = (dec >= UT[width](lo)) & (dec < UT[width](hi))
            r = D.query_filter_sum(col, A.PRED_RANGE, lo, hi).cpu().numpy()
            assert (int(r[0]), int(r[1])) == (int(sel.sum()), _sum_mod64(dec[sel])), (tag, lo, hi)


@pytest.mark.parametrize("width", [32, 64])
def test_compute_is_zero(width):
    # kernels.cuh:56-89: every value is &= 0 each repetition, so "== 1" never holds.
    rng = np.random.default_rng(13)
    packed, bws, offs, bases = random_column(rng, 16, width, [9] * 16)
    col = gpu_column(packed, bws, offs, bases, width)
    assert D.compute(col, 10).item() == 0


@pytest.mark.parametrize("width", [32, 64])
def test_multi_column(width):
    from flsgpu import encode as E
    rng = np.random.default_rng(14)
    n_vecs = 6
    value = 12345
    cols_hit, cols_miss = [], []
    for c in range(10):
        vals = rng.integers(0, 1 << 20, n_vecs * 1024).astype(UT[width])
        hit = vals.copy()
        v = 3
        lane = 5
        nl = 1024 // width
        hit[v * 1024 + lane + nl * np.arange(width)] = value  # one full lane equal in every column
        p, b, o, bs = E.ffor_encode_per_vector(hit)
        cols_hit.append(gpu_column(p, b, o, bs, width))
        miss = hit.copy()
        if c == 4:
            miss[v * 1024 + lane + nl * 7] = value + 1  # one column breaks the lane
        p, b, o, bs = E.ffor_encode_per_vector(miss)
        cols_miss.append(gpu_column(p, b, o, bs, width))
    for n in range(1, 11):
        assert D.query_multi_column(cols_hit[:n], value).item() == 1
        assert D.query_multi_column(cols_miss[:n], value).item() == (1 if n <= 4 else 0)
        assert D.query_multi_column(cols_hit[:n], value + 7).item() == 0
