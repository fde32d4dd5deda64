"""Product-side encoders and synthetic column generators (host numpy and device torch).

FastLanes packing follows the reference layout (src/fls/ffor.cpp, e.g.
ffor_3bit_32ow :2014): the value at logical index l + N_LANES*j goes to bits
[j*bw, (j+1)*bw) of lane l's word stream, word r of lane l at l + N_LANES*r.
Here it is vectorised over all vectors of a column at once.
"""
from __future__ import annotations

import numpy as np
import torch

_UT = {32: np.uint32, 64: np.uint64, 16: np.uint16, 8: np.uint8}


def n_lanes(width: int) -> int:
    return 1024 // width


def fls_pack(values: np.ndarray, bw: int) -> np.ndarray:
    """Pack (n_vecs*1024,) unsigned values (already FOR-subtracted) at width bw.
    Returns (n_vecs * bw * N_LANES,) words, vector v at offset v*bw*N_LANES."""
    ut = values.dtype.type
    w = values.dtype.itemsize * 8
    nl = n_lanes(w)
    v = values.reshape(-1, 1024)
    nv = v.shape[0]
    out = np.zeros((nv, bw * nl), dtype=ut)
    if bw == 0:
        return out.reshape(-1)
    mask = ut((1 << bw) - 1) if bw < w else ut(~ut(0))
    for j in range(w):
        x = v[:, nl * j: nl * (j + 1)] & mask
        p = j * bw
        r, s = divmod(p, w)
        out[:, r * nl:(r + 1) * nl] |= (x << ut(s)).astype(ut)
        if s + bw > w:
            out[:, (r + 1) * nl:(r + 2) * nl] |= (x >> ut(w - s)).astype(ut)
    return out.reshape(-1)


def ffor_encode_uniform(values: np.ndarray, bw: int | None = None):
    """FFOR with base = vector minimum and one bit width for the whole column
    (the widest vector's if bw is None). Returns packed, bit_widths,
    vector_offsets, bases."""
    ut = values.dtype.type
    w = values.dtype.itemsize * 8
    v = values.reshape(-1, 1024)
    bases = v.min(axis=1)
    deltas = (v - bases[:, None]).astype(ut)
    if bw is None:
        mx = int(deltas.max()) if deltas.size else 0
        bw = mx.bit_length()
    nv = v.shape[0]
    packed = fls_pack(deltas.reshape(-1), bw)
    bit_widths = np.full(nv, bw, dtype=np.uint8)
    offsets = np.arange(nv, dtype=np.uint64) * np.uint64(bw * n_lanes(w))
    return packed, bit_widths, offsets, bases.astype(ut)


def ffor_encode_per_vector(values: np.ndarray):
    """FFOR with base = vector minimum and each vector at its own minimal
    bit width (analyze_ffor, src/alp/encoder.hpp:104-115)."""
    ut = values.dtype.type
    w = values.dtype.itemsize * 8
    nl = n_lanes(w)
    v = values.reshape(-1, 1024)
    nv = v.shape[0]
    bases = v.min(axis=1)
    deltas = (v - bases[:, None]).astype(ut)
    mx = deltas.max(axis=1)
    bws = np.array([int(m).bit_length() for m in mx], dtype=np.uint8)
    offsets = np.zeros(nv, dtype=np.uint64)
    if nv:
        offsets[1:] = np.cumsum(bws.astype(np.uint64) * np.uint64(nl))[:-1]
    parts = []
    for b in np.unique(bws):
        idx = np.nonzero(bws == b)[0]
        parts.append((idx, fls_pack(deltas[idx].reshape(-1), int(b)).reshape(len(idx), -1)))
    total = int(bws.astype(np.int64).sum()) * nl
    packed = np.zeros(total, dtype=ut)
    for idx, words in parts:
        for k, vi in enumerate(idx):
            o = int(offsets[vi])
            packed[o:o + words.shape[1]] = words[k]
    return packed, bws, offsets, bases.astype(ut)


# ---------------------------------------------------------------------------
# Device-side synthetic columns (benchmarks): the packed words of a column
# whose offsets are uniform in [0, 2^bw) are uniformly random bits, so they
# are drawn directly on the GPU.
# ---------------------------------------------------------------------------
def random_words_device(n_words: int, width: int, gen: torch.Generator, device="cuda"):
    if n_words == 0:  # keep a valid device address for all-zero-width columns
        return torch.zeros(1, dtype=torch.uint32 if width == 32 else torch.uint64, device=device)
    n32 = n_words * (width // 32)
    t = torch.randint(-(2 ** 31), 2 ** 31, (n32,), dtype=torch.int32, device=device, generator=gen)
    return t.view(torch.uint32) if width == 32 else t.view(torch.uint64)


def synthetic_ffor_column_device(n_vecs: int, bw: int, width: int, seed: int, device="cuda",
                                 base_max: int | None = None):
    """Uniform-width FFOR column: per-vector random base in [0, base_max) plus
    uniform offsets in [0, 2^bw). Returns (packed, bit_widths, vector_offsets,
    bases) device tensors."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    nl = n_lanes(width)
    packed = random_words_device(n_vecs * bw * nl, width, gen, device)
    bit_widths = torch.full((n_vecs,), bw, dtype=torch.uint8, device=device)
    offsets = (torch.arange(n_vecs, dtype=torch.int64, device=device) * (bw * nl)).view(torch.uint64)
    if base_max is None:
        base_max = 2 ** 30
    bases = torch.randint(0, base_max, (n_vecs,), dtype=torch.int64, device=device, generator=gen)
    bases = bases.to(torch.int32).view(torch.uint32) if width == 32 else bases.view(torch.uint64)
    return packed, bit_widths, offsets, bases
