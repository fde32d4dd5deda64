"""Reductions on several streams at once: each (device, stream) pair has its
own partials area, so fused filter + sum and query-any launches in flight on
two streams both return their own column's answer. Also the fused filter +
sum and compute entry points for plain bit-packed (BP) columns."""
import numpy as np
import pytest

import oracle_py as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import device as D  # noqa: E402

M64 = (1 << 64) - 1


def column(rng, n_vecs, bw, width, with_bases=True):
    ut = np.uint32 if width == 32 else np.uint64
    nl = 1024 // width
    bws = np.full(n_vecs, bw, np.uint8)
    offs = np.arange(n_vecs, dtype=np.uint64) * np.uint64(bw * nl)
    packed = np.frombuffer(rng.bytes(max(n_vecs * bw * nl * width // 8, 8)), ut).copy()
    bases = (rng.integers(0, 1 << 20, n_vecs).astype(ut) if with_bases else None)
    dec = O.ffor_column_decode(O.FFORColumn(n_vecs * 1024, packed, bws, offs,
                                            bases if bases is not None else np.zeros(n_vecs, ut)))
    return D.DeviceFFORColumn.from_host(packed, bws, offs, bases, n_vecs * 1024), dec


def exact_sum(a):
    return int(a.astype(np.uint64).sum(dtype=np.uint64))  # numpy wraps mod 2^64


@pytest.mark.parametrize("width", [32, 64])
def test_two_streams_reduce_concurrently(width):
    rng = np.random.default_rng(70 + width)
    cols = [column(rng, 4096, bw, width) for bw in (9, 17)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    res = [torch.zeros(2, dtype=torch.uint64, device="cuda") for _ in cols]
    found = [torch.zeros(1, dtype=torch.uint8, device="cuda") for _ in cols]
    torch.cuda.synchronize()
    for rep in range(3):
        # interleave launches on the two streams so their reductions overlap
        for k in range(4):
            for i, ((c, _), s) in enumerate(zip(cols, streams)):
                D.query_filter_sum(c, A.PRED_ALL, result=res[i], stream=s)
                D.query_any(c, int(cols[i][1][123 + k]), result=found[i], stream=s)
        torch.cuda.synchronize()
        for i, (c, dec) in enumerate(cols):
            r = res[i].cpu().numpy()
            assert int(r[0]) == dec.size and int(r[1]) == exact_sum(dec), (rep, i)
            assert int(found[i].item()) == 1, (rep, i)
    L = A.lib()
    for s in streams:
        assert L.flsgpu_release_stream(C_stream(s)) == 0
    # a released stream gets a new area on its next reduction
    D.query_filter_sum(cols[0][0], A.PRED_ALL, result=res[0], stream=streams[0])
    torch.cuda.synchronize()
    assert int(res[0].cpu().numpy()[1]) == exact_sum(cols[0][1])


def C_stream(s):
    import ctypes as C
    return C.c_void_p(s.cuda_stream)


@pytest.mark.parametrize("width", [32, 64])
def test_bp_filter_sum_and_compute(width):
    rng = np.random.default_rng(90 + width)
    for bw in (0, 1, 13, width - 1, width):
        col, dec = column(rng, 37, bw, width, with_bases=False)
        assert col.bases is None
        r = D.query_filter_sum(col, A.PRED_ALL).cpu().numpy()
        assert int(r[0]) == dec.size and int(r[1]) == exact_sum(dec), bw
        r = D.query_filter_sum(col, A.PRED_OFFSET_LOWER_HALF).cpu().numpy()
        sel = dec.astype(np.uint64) < np.uint64(1 << (bw - 1)) if bw > 0 else np.zeros(dec.shape, bool)
        assert int(r[0]) == int(sel.sum()) and int(r[1]) == exact_sum(dec[sel]), bw
        lo = int(dec[5])
        hi = min(lo + (1 << max(bw - 2, 0)), (1 << width) - 1)
        r = D.query_filter_sum(col, A.PRED_RANGE, lo, hi).cpu().numpy()
        d = dec.astype(np.uint64)
        sel = (d >= np.uint64(lo)) & (d < np.uint64(hi))
        assert int(r[0]) == int(sel.sum()) and int(r[1]) == exact_sum(dec[sel]), bw
        assert int(D.compute(col, 3).item()) == 0
