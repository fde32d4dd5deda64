#!/usr/bin/env python3
"""Writes the golden fixtures in this directory from the reference's own CPU
codec, compiled from its sources into oracle/_ref (see oracle/Makefile).

Each fixture holds seeded inputs and the reference's outputs only (no
reference source). Regenerate with:  make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle_py as O  # noqa: E402

UT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def rand_bits(rng, n, w):
    return np.frombuffer(rng.bytes(n * w // 8), dtype=UT[w]).copy()


def fls_fixture(rng):
    out = {}
    for w in (8, 16, 32, 64):
        widths = sorted({0, 1, 3, 7, w // 2 + 1, w - 1, w} & set(range(w + 1)))
        for bw in widths:
            packed = rand_bits(rng, bw * (1024 // w), w)
            base = rand_bits(rng, 1, w)[0]
            out[f"u{w}_bw{bw}_packed"] = packed
            out[f"u{w}_bw{bw}_base"] = np.array([base], UT[w])
            out[f"u{w}_bw{bw}_unffor"] = O.unffor(packed, bw, base, use_ref=True)
            vals = rand_bits(rng, 1024, w)
            out[f"u{w}_bw{bw}_values"] = vals
            out[f"u{w}_bw{bw}_ffor"] = O.ffor(vals, bw, base, use_ref=True)
    np.savez_compressed(os.path.join(HERE, "fls_vectors.npz"), **out)


def alp_decode_fixture(rng):
    out = {}
    for w, ft, nf, ne in ((32, np.float32, 10, 11), (64, np.float64, 19, 21)):
        for bw in (0, 1, 5, 13, w // 2, w - 1):
            for k in range(3):
                packed = rand_bits(rng, bw * (1024 // w), w)
                base = int(rng.integers(0, 1 << 20))
                fac = int(rng.integers(0, nf))
                exp = int(rng.integers(0, ne))
                key = f"f{w}_bw{bw}_{k}"
                out[key + "_packed"] = packed
                out[key + "_params"] = np.array([base, fac, exp], np.uint64)
                out[key + "_out"] = O.alp_decode_vec(packed, bw, base, fac, exp, ft, use_ref=True)
    np.savez_compressed(os.path.join(HERE, "alp_decode_vectors.npz"), **out)


def datasets(rng):
    n = 8 * 1024
    d = {}
    k = rng.integers(-1_000_000, 1_000_001, n)
    x = k / 100.0
    x[rng.choice(n, 80, replace=False)] = rng.uniform(-1e4, 1e4, 80)
    d["f64_dec2_outliers"] = x
    d["f64_ints"] = rng.integers(-5000, 5000, n).astype(np.float64)
    d["f64_dec4"] = np.round(rng.uniform(0, 100, n), 4)
    y = np.round(rng.uniform(-50, 50, n), 1)
    y[17] = np.nan
    y[300] = np.inf
    y[4000] = -0.0
    d["f64_dec1_specials"] = y
    kf = rng.integers(-10_000, 10_001, n).astype(np.float32)
    d["f32_dec1"] = kf / np.float32(10.0)
    d["f32_ints"] = rng.integers(-3000, 3000, n).astype(np.float32)
    d["f32_dec2"] = (rng.integers(-2_000, 2_000, n) / 100).astype(np.float32)
    return d


def alp_encode_fixture(rng):
    out = {}
    for name, x in datasets(rng).items():
        col = O.alp_column_encode(x, use_ref=True)
        assert col is not None, name
        out[name + "_input"] = x
        for f in ("factor_indices", "fraction_indices", "counts", "positions", "exceptions"):
            out[name + "_" + f] = getattr(col, f)
        out[name + "_bit_widths"] = col.ffor.bit_widths
        out[name + "_bases"] = col.ffor.bases
        out[name + "_packed"] = col.ffor.packed
        out[name + "_top_k"] = np.array(O.alp_top_k(x, use_ref=True), np.uint8).reshape(-1, 2)
    np.savez_compressed(os.path.join(HERE, "alp_encode_columns.npz"), **out)


def alprd_fixture(rng):
    """Dictionary choice on inputs whose left-part counts have no ties."""
    out = {}
    for w, ft in ((64, np.float64), (32, np.float32)):
        for case in range(3):
            # exponents with strictly decreasing, distinct frequencies
            n = 288
            counts = [100, 60, 40, 30, 20, 15, 10, 8, 5][: 9 - case]
            s = sum(counts)
            counts[0] += n - s
            parts = []
            for i, c in enumerate(counts):
                parts.append(rng.uniform(2.0 ** (-i - 1), 2.0 ** (-i), c))
            sample = np.concatenate(parts).astype(ft)
            rng.shuffle(sample)
            rbw, d, lbw = O.alprd_find_dictionary(sample, use_ref=True)
            key = f"f{w}_{case}"
            out[key + "_sample"] = sample
            out[key + "_choice"] = np.array([rbw, lbw], np.uint32)
            out[key + "_dict"] = d
    np.savez_compressed(os.path.join(HERE, "alprd_dictionaries.npz"), **out)


def main():
    if O.ref() is None:
        sys.exit("oracle/_ref/libflsref.so is missing: run `make -C oracle` where the reference exists")
    rng = np.random.default_rng(20251205)
    fls_fixture(rng)
    alp_decode_fixture(rng)
    alp_encode_fixture(rng)
    alprd_fixture(rng)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
