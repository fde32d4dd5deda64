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
@pytest.mark.parametrize("variant", [A.VARIANT_LANE, A.VARIANT_LANE_NT])
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
