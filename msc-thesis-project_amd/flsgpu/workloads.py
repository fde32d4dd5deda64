"""Seeded synthetic workloads (numpy, host side) shared by the benchmark and
the GPU tests: the ALP decimal columns and the mixed-encoding batch."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _capi as A
from . import alp as AL


def alp_decimal_values(width: int, n: int, rng, decimals: int | None = None,
                       outliers_per_vector: int = 10):
    """Values with a fixed number of decimals (2^26 doubles k/100 with
    k ~ U[-10^6, 10^6], or floats k/10 with k ~ U[-10^4, 10^4] by default) and
    `outliers_per_vector` distinct positions per 1024-vector replaced by
    full-precision uniform values."""
    if width == 64:
        if decimals is None:
            x = rng.integers(-1_000_000, 1_000_001, n) / 100.0
        else:
            x = np.round(rng.uniform(-1e4, 1e4, n), decimals)
        repl = rng.uniform(-1e4, 1e4, (n // 1024, outliers_per_vector))
    else:
        x = rng.integers(-10_000, 10_001, n).astype(np.float32) / np.float32(10.0)
        repl = rng.uniform(-1e3, 1e3, (n // 1024, outliers_per_vector)).astype(np.float32)
    if outliers_per_vector:
        nv = n // 1024  # distinct positions: an odd stride is coprime with 1024
        pos = (rng.integers(0, 1024, (nv, 1)) + (2 * rng.integers(0, 512, (nv, 1)) + 1)
               * np.arange(outliers_per_vector)[None, :]) % 1024
        np.put_along_axis(x.reshape(-1, 1024), pos, repl, axis=1)
    return x


@dataclass
class BatchColumnSpec:
    kind: str          # "ffor_u32" | "alp_f64" | "alprd_f64" | ...
    host: object       # host column (tuple for FFOR, AlpHostColumn, AlprdHostColumn)
    values: np.ndarray  # the original values (decode must reproduce them bit for bit)


def ffor_u32_column(n: int, bw: int, rng):
    nv = n // 1024
    bases = rng.integers(0, 1 << 31, nv).astype(np.uint32)
    offs = rng.integers(0, 1 << bw, n, dtype=np.uint64).astype(np.uint32)
    values = (offs.reshape(nv, 1024) + bases[:, None]).reshape(-1).astype(np.uint32)
    bws = np.full(nv, bw, np.uint8)
    packed, voffs = AL.ffor_pack(values, bws, bases)
    return (packed, bws, voffs, bases), values


def mixed_batch(seed: int, n_columns: int = 96, min_log2: int = 12, max_log2: int = 24):
    """Columns with log-uniform lengths in [2^min_log2, 2^max_log2] (multiples
    of 1024): a third FFOR uint32 (one random width 1..24 per column), a third
    ALP double (0..4 decimals, 0..3% full-precision outliers, G-ALP layout),
    a third ALP-RD double (uniform in [0, 1))."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_columns):
        n = max(1024, int(round(2.0 ** rng.uniform(min_log2, max_log2) / 1024)) * 1024)
        kind = i % 3
        if kind == 0:
            host, values = ffor_u32_column(n, int(rng.integers(1, 25)), rng)
            specs.append(BatchColumnSpec("ffor_u32", host, values))
        elif kind == 1:
            d = int(rng.integers(0, 5))
            x = np.round(rng.uniform(-1e4, 1e4, n), d)
            p = rng.uniform(0.0, 0.03)
            k = int(round(p * n))
            if k:
                x[rng.choice(n, k, replace=False)] = rng.uniform(-1e4, 1e4, k)
            col = AL.alp_encode(x)
            if col is None:  # sampling chose ALP-RD for this column
                specs.append(BatchColumnSpec("alprd_f64", AL.alprd_encode(x), x))
            else:
                try:
                    specs.append(BatchColumnSpec("alp_ext_f64", AL.galp_convert(col), x))
                except A.FlsgpuError:  # a lane holds more than 63 exceptions
                    specs.append(BatchColumnSpec("alp_f64", col, x))
        else:
            x = rng.uniform(0.0, 1.0, n)
            specs.append(BatchColumnSpec("alprd_f64", AL.alprd_encode(x), x))
    return specs


def to_device(spec: BatchColumnSpec, device="cuda"):
    from . import device as D
    if spec.kind.startswith("ffor") or spec.kind.startswith("bp"):
        p, b, o, s = spec.host
        return D.DeviceFFORColumn.from_host(p, b, o, None if spec.kind.startswith("bp") else s,
                                            len(spec.values), device)
    if spec.kind.startswith("alprd"):
        return AL.DeviceAlprdColumn.from_host(spec.host, device)
    return AL.DeviceAlpColumn.from_host(spec.host, device)
