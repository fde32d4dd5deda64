"""Full-size checks at the benchmark configurations, through properties that
do not need the CPU oracle to run at that size:

- C1: 2^25 uint32 values 1e6 + U[0, 2^13) at width 13 decode back to the
  values, and the fused sum equals their sum (numpy);
- C2: 2^28-value columns of the bit-width sweep decode, with the autotuned
  variant, to what a plain PyTorch restatement of the FastLanes layout
  computes on the device (value l + N_LANES*j of a vector = bits
  [j*bw, (j+1)*bw) of lane l's words, plus the vector's base mod 2^W), and
  the fused filter + sum equals the PyTorch count and sum;
- C3: 2^26 doubles (2 decimals, 1% exceptions) and 2^27 floats round-trip
  bit for bit through the host encoder and the GPU decoder, in both
  exception layouts.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import alp as AL  # noqa: E402
from flsgpu import autotune as AT  # noqa: E402
from flsgpu import device as D  # noqa: E402
from flsgpu import encode as E  # noqa: E402

MASK64 = (1 << 64) - 1


@pytest.fixture(scope="module", autouse=True)
def tuned():
    A.check(A.lib().flsgpu_init(0))
    AT.load_table()
    yield
    torch.cuda.empty_cache()


def lsr(x, s):
    """Logical right shift of int64 words (torch shifts int64 arithmetically)."""
    if s == 0:
        return x
    return (x >> s) & ((1 << (64 - s)) - 1)


def torch_ffor_decode(packed, bases, n_vecs, bw, width):
    """Uniform-width FFOR decode restated with PyTorch int64 ops on the device;
    returns int64 values (uint32 values zero-extended, uint64 as their bits)."""
    nl = 1024 // width
    words = packed.view(torch.int32).to(torch.int64) & 0xFFFFFFFF if width == 32 \
        else packed.view(torch.int64)
    w = words.view(n_vecs, bw, nl) if bw else None
    base = bases.view(torch.int32).to(torch.int64) & 0xFFFFFFFF if width == 32 \
        else bases.view(torch.int64)
    out = torch.empty(n_vecs, width, nl, dtype=torch.int64, device=packed.device)
    for j in range(width):
        if bw == 0:
            v = torch.zeros(n_vecs, nl, dtype=torch.int64, device=packed.device)
        else:
            r, s = divmod(j * bw, width)
            lo = w[:, r, :]
            v = lsr(lo, s) if width == 64 else (lo >> s)
            if s + bw > width:
                hi = w[:, r + 1, :]
                v = v | (hi << (width - s))
            if bw < width:
                v = v & ((1 << bw) - 1)
            elif width == 32:
                v = v & 0xFFFFFFFF
        v = v + base[:, None]
        out[:, j, :] = v & 0xFFFFFFFF if width == 32 else v
    return out.view(-1)


def test_c1_decode_and_sum_2p25():
    rng = np.random.default_rng(2025)
    vals = (1_000_000 + rng.integers(0, 1 << 13, 1 << 25)).astype(np.uint32)
    p, b, o, s = E.ffor_encode_uniform(vals, 13)
    col = D.DeviceFFORColumn.from_host(p, b, o, s, len(vals))
    got = D.decompress(col).view(torch.int32).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, vals)
    r = D.query_filter_sum(col, A.PRED_ALL).cpu().numpy()
    assert int(r[0]) == len(vals) and int(r[1]) == int(vals.astype(np.uint64).sum())


@pytest.mark.parametrize("width,bw", [(32, 0), (32, 3), (32, 13), (32, 32), (64, 7), (64, 33),
                                      (64, 64)])
def test_c2_sweep_column_2p28(width, bw):
    n = 1 << 28
    n_vecs = n // 1024
    packed, bws, offs, bases = E.synthetic_ffor_column_device(n_vecs, bw, width, seed=width * 100 + bw)
    col = D.DeviceFFORColumn(n, packed, bws, offs, bases, width)
    ref = torch_ffor_decode(packed, bases, n_vecs, bw, width)
    got = D.decompress(col)
    got64 = got.view(torch.int32).to(torch.int64) & 0xFFFFFFFF if width == 32 else got.view(torch.int64)
    assert torch.equal(got64, ref)
    del got, got64
    # fused filter (offset < 2^(bw-1)) + sum against the same restatement
    r = D.query_filter_sum(col, A.PRED_OFFSET_LOWER_HALF).cpu().numpy()
    base = (bases.view(torch.int32).to(torch.int64) & 0xFFFFFFFF if width == 32
            else bases.view(torch.int64)).repeat_interleave(1024)
    off = ref - base
    if width == 32:
        off = off & 0xFFFFFFFF
    if bw == 0:
        sel = torch.zeros_like(off, dtype=torch.bool)
    elif bw == 64:
        sel = off >= 0  # top bit of the 64-bit offset clear
    else:
        sel = (off >= 0) & (off < (1 << (bw - 1)))
    count = int(sel.sum().item())
    # the sum mod 2^64 from 2^28 values: split into 32-bit halves to stay exact
    lo = (ref & 0xFFFFFFFF)[sel].sum().item()
    hi = lsr(ref, 32)[sel].sum().item()
    total = (int(lo) + (int(hi) << 32)) & MASK64
    assert int(r[0]) == count and int(r[1]) == total


@pytest.mark.parametrize("width", [64, 32])
def test_c3_alp_roundtrip_full_size(width):
    import bench as BN  # the benchmark's own C3 generator
    rng = np.random.default_rng(33)
    x = BN.c3_values(np, width, rng)
    col = AL.alp_encode(x)
    assert col is not None
    ut = np.uint64 if width == 64 else np.uint32
    for host in (col, AL.galp_convert(col)):
        d = AL.DeviceAlpColumn.from_host(host)
        got = AL.decompress(d).cpu().numpy()
        assert np.array_equal(got.view(ut), x.view(ut))
        del d
        torch.cuda.empty_cache()
