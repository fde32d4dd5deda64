"""The library's host encoder (include/flsgpu_host.h) against the oracle
encoder, which is itself pinned to the reference (test_oracle.py)."""
import numpy as np
import pytest

import oracle_py as O
from flsgpu import alp as AL
from flsgpu import encode as E

FT = {32: np.float32, 64: np.float64}


def bits(a):
    return a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def datasets(rng, n):
    out = []
    x = rng.integers(-1_000_000, 1_000_001, n) / 100.0
    x[rng.choice(n, n // 100, replace=False)] = rng.uniform(-1e4, 1e4, n // 100)
    out.append(x)
    out.append(np.round(rng.uniform(-1e3, 1e3, n), int(rng.integers(0, 5))))
    y = np.round(rng.uniform(0, 50, n), 1)
    y[[1, 1500, 7000]] = [np.nan, -np.inf, -0.0]
    out.append(y)
    out.append((rng.integers(-10_000, 10_001, n).astype(np.float32) / np.float32(10)))
    out.append(rng.integers(-3000, 3000, n).astype(np.float32))
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_alp_encoder_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    for x in datasets(rng, 1024 * 110):
        a = AL.alp_encode(x, n_threads=4)
        b = O.alp_column_encode(x)
        assert (a is None) == (b is None)
        if a is None:
            continue
        for f in ("factor_indices", "fraction_indices", "counts", "positions", "exceptions_offsets"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert np.array_equal(bits(a.exceptions), bits(b.exceptions))
        assert np.array_equal(a.packed, b.ffor.packed)
        assert np.array_equal(a.bit_widths, b.ffor.bit_widths)
        assert np.array_equal(a.bases, b.ffor.bases)
        assert np.array_equal(a.vector_offsets, b.ffor.vector_offsets)
        assert a.compressed_bytes_alp == b.compressed_bytes_alp
        assert a.compressed_bytes_galp == b.compressed_bytes_galp
        g1 = AL.galp_convert(a)
        g2 = O.galp_convert(b)
        assert np.array_equal(g1.offsets_counts, g2.offsets_counts)
        assert np.array_equal(g1.positions, g2.positions)
        assert np.array_equal(bits(g1.exceptions), bits(g2.exceptions))


def test_alp_encoder_scheme_switch_and_search():
    rng = np.random.default_rng(3)
    assert AL.alp_encode(rng.uniform(0, 1, 1024 * 20)) is None  # ALP_RD data
    for ft in (np.float32, np.float64):
        samples = np.round(rng.uniform(-100, 100, (50, 32)), rng.integers(0, 4)).astype(ft)
        samples[3, :] = rng.uniform(-1, 1, 32)
        e, f, est = AL.alp_search(samples, n_threads=2)
        for v in range(50):
            oe, of, oest = O.alp_search_one(samples[v])
            assert (e[v], f[v], est[v]) == (oe, of, oest), v


def test_encode_with_params_roundtrip():
    rng = np.random.default_rng(4)
    x = np.round(rng.uniform(-10, 10, 1024 * 8), 3)
    exps = rng.integers(0, 19, 8).astype(np.uint8)
    facs = np.array([rng.integers(0, e + 1) for e in exps], np.uint8)
    col = AL.alp_encode_with_params(x, exps, facs)
    ocol = O.ALPColumn(col.n_values, O.FFORColumn(col.n_values, col.packed, col.bit_widths,
                                                  col.vector_offsets, col.bases),
                       col.factor_indices, col.fraction_indices, col.exceptions_offsets,
                       col.exceptions if len(col.exceptions) else np.zeros(1),
                       col.positions if len(col.positions) else np.zeros(1, np.uint16), col.counts)
    assert np.array_equal(bits(O.alp_column_decode(ocol)), bits(x))
    with pytest.raises(Exception):
        AL.alp_encode_with_params(x, np.full(8, 19, np.uint8), facs)


def test_galp_rejects_64_exceptions_in_a_lane():
    n_vecs = 1
    pos = (3 + 16 * np.arange(64)).astype(np.uint16)
    col = AL.AlpHostColumn(1024, np.zeros(1, np.uint64), np.zeros(1, np.uint8), np.zeros(1, np.uint64),
                           np.zeros(1, np.uint64), np.zeros(1, np.uint8), np.zeros(1, np.uint8),
                           np.zeros(1, np.uint64), np.ones(64), pos, np.array([64], np.uint16))
    with pytest.raises(Exception, match="63"):
        AL.galp_convert(col)
    del n_vecs


@pytest.mark.parametrize("w", [32, 64])
def test_alprd_encoder_matches_oracle(w):
    rng = np.random.default_rng(10 + w)
    for x in (rng.uniform(0, 1, 1024 * 120), rng.normal(0, 1e5, 1024 * 30)):
        x = x.astype(FT[w])
        a = AL.alprd_encode(x)
        b = O.alprd_column_encode(x)
        assert (a.right_bw, a.left_bw) == (b.right_bw, b.left_bw)
        assert np.array_equal(a.dictionary, b.dictionary)
        assert np.array_equal(a.right_packed, b.right_packed[: len(a.right_packed)])
        assert np.array_equal(a.left_packed, b.left_packed[: len(a.left_packed)])
        assert np.array_equal(a.counts, b.counts) and np.array_equal(a.positions, b.positions)
        assert np.array_equal(a.exceptions, b.exceptions)


@pytest.mark.parametrize("w", [32, 64])
def test_ffor_pack_matches_oracle(w):
    rng = np.random.default_rng(20 + w)
    ut = np.uint32 if w == 32 else np.uint64
    vals = np.frombuffer(rng.bytes(1024 * 6 * w // 8), dtype=ut).copy()
    for bw in (0, 1, 13, w - 1, w):
        p, b, o, s = E.ffor_encode_uniform(vals & ut((1 << bw) - 1 if bw < w else (1 << w) - 1), bw)
        for v in range(6):
            seg = p[v * bw * (1024 // w):(v + 1) * bw * (1024 // w)]
            assert np.array_equal(O.unffor(seg, bw, s[v]), (vals[v * 1024:(v + 1) * 1024] &
                                  ut((1 << bw) - 1 if bw < w else (1 << w) - 1)))
    p2, b2, o2, s2 = E.ffor_encode_per_vector(vals >> ut(w // 2))
    col = O.FFORColumn(len(vals), p2, b2, o2, s2)
    assert np.array_equal(O.ffor_column_decode(col), vals >> ut(w // 2))
    lib_p, lib_o = AL.ffor_pack(vals >> ut(w // 2), b2, s2)
    assert np.array_equal(lib_p, p2) and np.array_equal(lib_o, o2)
