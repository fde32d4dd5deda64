"""GPU parity of the ALP-RD decoder: bitwise equal to the oracle's decode of
the same encoded buffers and to the original values (the reference pins
ALP-RD only through round trips: rd.hpp:103-172 has no caller or test)."""
import numpy as np
import pytest

import oracle_py as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import alp as AL  # noqa: E402

FT = {32: np.float32, 64: np.float64}


def bits(a):
    return a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def to_oracle(c):
    return O.ALPRDColumn(c.n_values, c.right_bw, c.left_bw, c.dictionary, c.right_packed,
                         c.left_packed, c.exceptions_offsets, c.exceptions, c.positions, c.counts)


@pytest.mark.parametrize("w", [32, 64])
@pytest.mark.parametrize("variant", [A.VARIANT_LANE, A.VARIANT_LANE_NT,
                                     A.VARIANT_HALF_LANE_STAGED, A.VARIANT_QUARTER_LANE_STAGED,
                                     A.VARIANT_EIGHTH_LANE_STAGED])
def test_alprd_roundtrip(w, variant):
    rng = np.random.default_rng(40 + w)
    cases = [rng.uniform(0, 1, 1024 * 37), rng.normal(0, 1e4, 1024 * 9),
             np.frombuffer(rng.bytes(1024 * 5 * w // 8), dtype=FT[w]).copy(),
             np.concatenate([rng.uniform(0, 1, 1024 * 3), rng.uniform(-1e9, 1e9, 1024 * 2)])]
    for x in cases:
        x = x.astype(FT[w])
        col = AL.alprd_encode(x)
        got = AL.alprd_decompress(AL.DeviceAlprdColumn.from_host(col), variant=variant).cpu().numpy()
        assert np.array_equal(bits(got), bits(x))
        assert np.array_equal(bits(got), bits(O.alprd_column_decode(to_oracle(col))))


def test_alprd_rejects_bad_widths():
    rng = np.random.default_rng(1)
    col = AL.alprd_encode(rng.uniform(0, 1, 2048))
    d = AL.DeviceAlprdColumn.from_host(col)
    d.right_bw = 40
    with pytest.raises(A.FlsgpuError):
        AL.alprd_decompress(d)


def random_rd_column(rng, w, n_vecs, right_bw, left_bw, max_exc):
    """Random parts at any right width: every left index is in the dictionary
    (dictionary_size = 2^left_bw), exceptions at distinct positions."""
    nl = 1024 // w
    ut = np.uint32 if w == 32 else np.uint64
    right = np.frombuffer(rng.bytes(n_vecs * right_bw * nl * w // 8), ut).copy()
    left = np.frombuffer(rng.bytes(n_vecs * left_bw * 64 * 2), np.uint16).copy()
    dictionary = rng.integers(0, 1 << (w - right_bw), 1 << left_bw).astype(np.uint16)
    counts = rng.integers(0, max_exc + 1, n_vecs).astype(np.uint16)
    eo = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)
    pos = np.concatenate([rng.choice(1024, int(c), replace=False) for c in counts]).astype(np.uint16)
    exc = rng.integers(0, 1 << (w - right_bw), len(pos)).astype(np.uint16)
    return AL.AlprdHostColumn(n_vecs * 1024, right_bw, left_bw, dictionary, right, left, eo, exc,
                              pos, counts)


@pytest.mark.parametrize("w", [32, 64])
def test_alprd_every_right_width(w):
    """All 16 right widths and left widths 1..3, every decoder variant, up to
    200 exceptions in a vector (more than one pass of a wave), ragged tails."""
    rng = np.random.default_rng(70 + w)
    for right_bw in range(w - 16, w):
        lbw = 1 + right_bw % 3
        col = random_rd_column(rng, w, 5 + right_bw % 4, right_bw, lbw, 200)
        exp = O.alprd_column_decode(to_oracle(col))
        d = AL.DeviceAlprdColumn.from_host(col)
        for variant in (A.VARIANT_LANE, A.VARIANT_HALF_LANE_STAGED, A.VARIANT_EIGHTH_LANE_STAGED):
            got = AL.alprd_decompress(d, variant=variant).cpu().numpy()
            assert np.array_equal(bits(got), bits(exp)), (right_bw, variant)


@pytest.mark.parametrize("w", [32, 64])
def test_alprd_staged_misaligned_output(w):
    rng = np.random.default_rng(80 + w)
    x = rng.uniform(0, 1, 1024 * 6).astype(FT[w])
    d = AL.DeviceAlprdColumn.from_host(AL.alprd_encode(x))
    buf = torch.zeros(len(x) + 1, dtype=torch.float32 if w == 32 else torch.float64, device="cuda")
    out = buf[1:]
    assert out.data_ptr() % 16 != 0
    AL.alprd_decompress(d, out, variant=A.VARIANT_QUARTER_LANE_STAGED)
    assert np.array_equal(bits(out.cpu().numpy()), bits(x)) and buf[0].item() == 0.0
