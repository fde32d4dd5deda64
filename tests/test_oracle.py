"""Pins the CPU oracle (oracle/*.c) before anything is checked against it:
golden fixtures written by the compiled reference (tests/golden), the
compiled reference itself where it was built here (oracle/_ref), and checks
fixed by the mathematics (closed forms, brute-force bit models, round trips,
invariants)."""
import os

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}
FT = {32: np.float32, 64: np.float64}
needs_ref = pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (no reference checkout)")


def rand_bits(rng, n, w):
    return np.frombuffer(rng.bytes(n * w // 8), dtype=UT[w]).copy()


def bits(a):
    return a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


# ---------------------------------------------------------------------------
# FastLanes
# ---------------------------------------------------------------------------
def test_fls_golden_vectors():
    g = np.load(os.path.join(GOLDEN, "fls_vectors.npz"))
    n = 0
    for key in g.files:
        if not key.endswith("_unffor"):
            continue
        pre = key[: -len("_unffor")]
        w = int(pre.split("_")[0][1:])
        bw = int(pre.split("_bw")[1])
        base = g[pre + "_base"][0]
        assert np.array_equal(O.unffor(g[pre + "_packed"], bw, base), g[key]), pre
        assert np.array_equal(O.ffor(g[pre + "_values"], bw, base), g[pre + "_ffor"]), pre
        n += 1
    assert n >= 25


def bitmodel_unpack(packed, bw, w):
    """Brute-force bit model: lane l's stream is the concatenation of words
    in[l + N*r] (LSB first); slot j is stream bits [j*bw, (j+1)*bw)."""
    nl = 1024 // w
    out = np.zeros(1024, dtype=object)
    for lane in range(nl):
        stream = 0
        for r in range(bw):
            stream |= int(packed[lane + nl * r]) << (w * r)
        for j in range(w):
            out[lane + nl * j] = (stream >> (j * bw)) & ((1 << bw) - 1)
    return out


@pytest.mark.parametrize("w", [8, 16, 32, 64])
def test_fls_bit_model_and_closed_forms(w):
    rng = np.random.default_rng(w)
    for bw in (0, 1, 2, 5, w // 2 - 1, w // 2 + 3, w - 1, w):
        if bw < 0 or bw > w:
            continue
        packed = rand_bits(rng, bw * (1024 // w), w)
        got = O.unpack(packed, bw)
        if bw:
            assert [int(x) for x in got] == list(bitmodel_unpack(packed, bw, w)), f"u{w} bw{bw}"
        base = rand_bits(rng, 1, w)[0]
        out = O.unffor(packed, bw, base)
        if bw == 0:
            assert (out == base).all()
        if bw == w:  # out = words, in lane-major order, + base (mod 2^W)
            nl = 1024 // w
            exp = np.array([(int(packed[l + nl * j]) + int(base)) % (1 << w)
                            for j in range(w) for l in range(nl)], dtype=object)
            assert [int(x) for x in out] == list(exp)
        # pack o unpack identity and ffor o unffor identity with wrap-around base
        vals = (rand_bits(rng, 1024, w) & UT[w]((1 << bw) - 1 if bw < w else (1 << w) - 1))
        assert np.array_equal(O.unpack(O.pack(vals, bw), bw), vals)
        shifted = (vals + base).astype(UT[w])
        assert np.array_equal(O.unffor(O.ffor(shifted, bw, base), bw, base), shifted)


@needs_ref
@pytest.mark.parametrize("w", [8, 16, 32, 64])
def test_fls_matches_compiled_reference_every_width(w):
    rng = np.random.default_rng(100 + w)
    for bw in range(w + 1):
        packed = rand_bits(rng, bw * (1024 // w), w)
        base = rand_bits(rng, 1, w)[0]
        assert np.array_equal(O.unffor(packed, bw, base), O.unffor(packed, bw, base, use_ref=True))
        assert np.array_equal(O.unpack(packed, bw), O.unpack(packed, bw, use_ref=True))
        vals = rand_bits(rng, 1024, w)
        assert np.array_equal(O.ffor(vals, bw, base), O.ffor(vals, bw, base, use_ref=True))
        assert np.array_equal(O.pack(vals, bw), O.pack(vals, bw, use_ref=True))


def test_ffor_column_decode_follows_offsets():
    rng = np.random.default_rng(5)
    bws = np.array([3, 0, 32, 17], np.uint8)
    offs = np.concatenate([[0], np.cumsum(bws.astype(np.uint64) * 32)[:-1]]).astype(np.uint64)
    packed = rand_bits(rng, int(bws.sum()) * 32, 32)
    bases = rand_bits(rng, 4, 32)
    col = O.FFORColumn(4096, packed, bws, offs, bases)
    out = O.ffor_column_decode(col)
    for v in range(4):
        seg = packed[int(offs[v]): int(offs[v]) + int(bws[v]) * 32]
        assert np.array_equal(out[v * 1024:(v + 1) * 1024], O.unffor(seg, int(bws[v]), bases[v]))


# ---------------------------------------------------------------------------
# ALP decode
# ---------------------------------------------------------------------------
FACT = {32: [10 ** i for i in range(10)], 64: [10 ** i for i in range(19)]}


def alp_model(packed, bw, base, fac, exp, w):
    """Integer part with exact Python integers, then one RNE int->float
    conversion and one IEEE multiply by the rounded 10^-exp table entry."""
    ft = FT[w]
    frac = ft(float(f"1e-{exp}")) if exp else ft(1.0)
    u = O.unpack(packed, bw)
    out = np.empty(1024, ft)
    for i in range(1024):
        x = ((int(u[i]) + int(base)) * FACT[w][fac]) % (1 << w)
        if x >= 1 << (w - 1):
            x -= 1 << w
        out[i] = ft(x) * frac  # numpy: int -> float RNE, float multiply
    return out


@pytest.mark.parametrize("w", [32, 64])
def test_alp_decode_integer_model(w):
    rng = np.random.default_rng(200 + w)
    nf, ne = (10, 11) if w == 32 else (19, 21)
    for bw in (0, 1, 9, 21, w - 1, w):
        packed = rand_bits(rng, bw * (1024 // w), w)
        base = int(rng.integers(0, 1 << 62)) % (1 << w)
        fac = int(rng.integers(0, nf))
        exp = int(rng.integers(0, ne))
        got = O.alp_decode_vec(packed, bw, base, fac, exp, FT[w])
        assert np.array_equal(bits(got), bits(alp_model(packed, bw, base, fac, exp, w))), (bw, fac, exp)


def test_alp_decode_golden_vectors():
    g = np.load(os.path.join(GOLDEN, "alp_decode_vectors.npz"))
    n = 0
    for key in g.files:
        if not key.endswith("_out"):
            continue
        pre = key[:-4]
        w = int(pre.split("_")[0][1:])
        bw = int(pre.split("_bw")[1].split("_")[0])
        base, fac, exp = (int(x) for x in g[pre + "_params"])
        got = O.alp_decode_vec(g[pre + "_packed"], bw, base, fac, exp, FT[w])
        assert np.array_equal(bits(got), bits(g[key])), pre
        n += 1
    assert n == 36


@needs_ref
@pytest.mark.parametrize("w", [32, 64])
def test_alp_decode_matches_compiled_reference(w):
    """Every width below W, every table entry pair for a few vectors. At
    width W the reference's generated falp multiplies the raw register by the
    fraction (falp.cpp:42198, :33293); the oracle keeps the decoder-API formula."""
    rng = np.random.default_rng(300 + w)
    nf, ne = (10, 11) if w == 32 else (19, 21)
    for bw in range(w):
        for _ in range(2):
            packed = rand_bits(rng, bw * (1024 // w), w)
            base = int(rng.integers(0, 1 << 62)) % (1 << w)
            fac = int(rng.integers(0, nf))
            exp = int(rng.integers(0, ne))
            a = O.alp_decode_vec(packed, bw, base, fac, exp, FT[w])
            b = O.alp_decode_vec(packed, bw, base, fac, exp, FT[w], use_ref=True)
            assert np.array_equal(bits(a), bits(b)), (bw, fac, exp)


# ---------------------------------------------------------------------------
# ALP encoder
# ---------------------------------------------------------------------------
def test_alp_encode_golden_columns():
    g = np.load(os.path.join(GOLDEN, "alp_encode_columns.npz"))
    names = sorted({k[: -len("_input")] for k in g.files if k.endswith("_input")})
    assert len(names) == 7
    for name in names:
        x = g[name + "_input"]
        col = O.alp_column_encode(x)
        assert col is not None
        assert O.alp_top_k(O.alp_sample(x)) == [tuple(int(v) for v in r) for r in g[name + "_top_k"]]
        for f in ("factor_indices", "fraction_indices", "counts", "positions"):
            assert np.array_equal(getattr(col, f), g[name + "_" + f]), (name, f)
        assert np.array_equal(bits(col.exceptions), bits(g[name + "_exceptions"])), name
        assert np.array_equal(col.ffor.bit_widths, g[name + "_bit_widths"]), name
        assert np.array_equal(col.ffor.bases, g[name + "_bases"]), name
        assert np.array_equal(col.ffor.packed, g[name + "_packed"]), name
        dec = O.alp_column_decode(col)
        assert np.array_equal(bits(dec), bits(x)), name  # lossless


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_alp_encode_matches_compiled_reference(seed):
    rng = np.random.default_rng(400 + seed)
    n = 1024 * 130  # more than one rowgroup
    kinds = [
        lambda: np.round(rng.uniform(-1e3, 1e3, n), int(rng.integers(0, 5))),
        lambda: rng.integers(0, 1 << 20, n).astype(np.float64) / 8.0,
        lambda: np.round(rng.normal(0, 50, n), 2).astype(np.float32),
        lambda: (rng.integers(-10_000, 10_000, n).astype(np.float32) / np.float32(10)),
    ]
    x = kinds[seed % len(kinds)]()
    x[rng.choice(n, 200, replace=False)] = rng.uniform(-1, 1, 200).astype(x.dtype)
    a = O.alp_column_encode(x)
    b = O.alp_column_encode(x, use_ref=True)
    assert (a is None) == (b is None)
    if a is None:
        return
    for f in ("factor_indices", "fraction_indices", "counts", "positions", "exceptions_offsets"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(bits(a.exceptions), bits(b.exceptions))
    for f in ("packed", "bit_widths", "bases", "vector_offsets"):
        assert np.array_equal(getattr(a.ffor, f), getattr(b.ffor, f)), f


@needs_ref
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_alp_column_decode_ranges(dtype):
    """The compiled reference's column decode (falp + patch_exceptions per
    vector, the CPU baseline's ALP leg) returns the encoded values, over the
    whole column and over split vector ranges."""
    from flsgpu import alp as AL
    rng = np.random.default_rng(77)
    n = 1024 * 37
    if dtype == np.float64:
        x = rng.integers(-10 ** 6, 10 ** 6, n) / 100.0
    else:
        x = rng.integers(-10 ** 4, 10 ** 4, n).astype(np.float32) / np.float32(10)
    x[rng.choice(n, 400, replace=False)] = rng.uniform(-1e3, 1e3, 400).astype(dtype)
    col = AL.alp_encode(x)
    assert len(col.exceptions) > 0
    whole = np.zeros(n, dtype=dtype)
    O.ref_alp_column_decode_range(col, 0, col.bit_widths.size, whole)
    assert np.array_equal(bits(whole), bits(x))
    parts = np.full(n, np.nan, dtype=dtype)
    for v0, v1 in ((0, 5), (5, 6), (6, 37)):
        O.ref_alp_column_decode_range(col, v0, v1, parts)
    assert np.array_equal(bits(parts), bits(x))


@needs_ref
def test_alp_encode_pieces_match_compiled_reference():
    rng = np.random.default_rng(7)
    for ft in (np.float32, np.float64):
        vec = np.round(rng.uniform(-100, 100, 1024), 2).astype(ft)
        vec[[3, 500, 1023]] = [np.nan, -0.0, np.inf]
        for fac, exp in ((0, 2), (1, 3), (0, 0), (2, 2)):
            e1, x1, p1 = O.alp_encode_vec(vec, fac, exp)
            e2, x2, p2 = O.alp_encode_vec(vec, fac, exp, use_ref=True)
            assert np.array_equal(e1, e2) and np.array_equal(bits(x1), bits(x2)) and np.array_equal(p1, p2)
            assert O.alp_analyze_ffor(e1) == O.alp_analyze_ffor(e1, use_ref=True)
        combos = [(2, 0), (3, 1), (4, 2), (1, 0), (5, 5)]
        assert O.alp_best_combination(combos, vec) == O.alp_best_combination(combos, vec, use_ref=True)
        for v in (1.5, -0.0, np.inf, 1e30, 123.456):
            assert O.alp_encode_value(ft(v), 1, 3) == O.alp_encode_value(ft(v), 1, 3, use_ref=True)
    x = np.round(rng.uniform(0, 10, 1024 * 40), 1)
    assert np.array_equal(O.alp_sample(x), O.alp_sample(x, use_ref=True))


def test_alp_fill_value_edge():
    """encode_simdized fills exception slots with the encoded value of the
    first non-exception slot; when slots 0..1022 are all exceptions the fill
    stays 0 (encoder.hpp:352-358)."""
    vec = np.full(1024, np.nan)
    vec[1023] = 5.0
    enc, exc, pos = O.alp_encode_vec(vec, 0, 0)
    assert len(pos) == 1023 and enc[1023] == 5 and (enc[:1023] == 0).all()
    vec = np.full(1024, np.nan)
    vec[700] = 9.0
    enc, exc, pos = O.alp_encode_vec(vec, 0, 0)
    assert (enc == 9).all()


# ---------------------------------------------------------------------------
# G-ALP
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("w", [32, 64])
def test_galp_invariants(w):
    rng = np.random.default_rng(600 + w)
    x = np.round(rng.uniform(-100, 100, 1024 * 12), 2).astype(FT[w])
    x[rng.choice(len(x), 300, replace=False)] = rng.uniform(-1, 1, 300).astype(FT[w])
    col = O.alp_column_encode(x)
    g = O.galp_convert(col)
    nl = 1024 // w
    for v in range(col.n_vecs):
        o, c = int(col.exceptions_offsets[v]), int(col.counts[v])
        pos, exc = col.positions[o:o + c], col.exceptions[o:o + c]
        order = np.argsort(pos % nl, kind="stable")  # stable partition by lane
        assert np.array_equal(g.positions[o:o + c], pos[order])
        assert np.array_equal(bits(g.exceptions[o:o + c]), bits(exc[order]))
        hdr = g.offsets_counts[v * nl:(v + 1) * nl]
        counts = np.bincount(pos % nl, minlength=nl)
        assert np.array_equal(hdr >> 10, counts)
        assert np.array_equal(hdr & 0x3FF, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    assert np.array_equal(bits(O.alp_column_decode(g)), bits(O.alp_column_decode(col)))


# ---------------------------------------------------------------------------
# ALP-RD
# ---------------------------------------------------------------------------
def test_alprd_golden_dictionaries():
    g = np.load(os.path.join(GOLDEN, "alprd_dictionaries.npz"))
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files})
    assert len(keys) == 6
    for k in keys:
        rbw, d, lbw = O.alprd_find_dictionary(g[k + "_sample"])
        assert [rbw, lbw] == list(g[k + "_choice"]) and np.array_equal(d, g[k + "_dict"]), k


@pytest.mark.parametrize("w", [32, 64])
def test_alprd_roundtrip(w):
    rng = np.random.default_rng(700 + w)
    for x in (rng.uniform(0, 1, 1024 * 30), rng.normal(0, 1e6, 1024 * 20),
              np.frombuffer(rng.bytes(1024 * 8 * w // 8), dtype=FT[w]).copy()):
        x = x.astype(FT[w])
        col = O.alprd_column_encode(x)
        assert w - 16 <= col.right_bw <= w - 1
        assert 1 <= col.left_bw <= 3 and 1 <= len(col.dictionary) <= 8
        assert np.array_equal(bits(O.alprd_column_decode(col)), bits(x))


@needs_ref
def test_alprd_decode_matches_reference_vector_decode():
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 1, 2048)
    col = O.alprd_column_encode(x)
    out = np.zeros(1024)
    right = O.unpack(col.right_packed[: col.right_bw * 16], col.right_bw)
    left = O.unpack(col.left_packed[: col.left_bw * 64], col.left_bw)
    c = int(col.counts[0])
    exc = col.exceptions[:c].copy()
    pos = col.positions[:c].copy()
    d = np.zeros(8, np.uint16)
    d[: len(col.dictionary)] = col.dictionary
    O.ref().ref_alprd_decode_vec_f64(out.ctypes.data, right.ctypes.data, left.ctypes.data,
                                     exc.ctypes.data, pos.ctypes.data, c, d.ctypes.data,
                                     len(col.dictionary), col.right_bw)
    assert np.array_equal(bits(out), bits(x[:1024]))
