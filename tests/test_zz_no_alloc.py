"""The no-allocation fused filter + sum
(flsgpu_query_filter_sum_ffor_u32_noalloc), called through the C ABI: its
[count, sum] equals flsgpu_query_filter_sum_ffor_u32's on the same column,
for every predicate, on two calls in a row (the second on a smaller column,
so stale partials of the first would show), and both equal the CPU oracle's
closed forms."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from flsgpu import _capi as A  # noqa: E402
from flsgpu import device as D  # noqa: E402


def column(rng, n_vecs, bw):
    nl = 32
    bws = np.full(n_vecs, bw, np.uint8)
    bws[::5] = max(bw - 3, 0)  # mixed widths in the column
    offs = np.zeros(n_vecs, np.uint64)
    offs[1:] = np.cumsum(bws.astype(np.uint64) * np.uint64(nl))[:-1]
    words = int(bws.astype(np.int64).sum()) * nl
    packed = np.frombuffer(rng.bytes(max(words, 1) * 4), np.uint32).copy()
    bases = rng.integers(0, 1 << 30, n_vecs).astype(np.uint32)
    dec = O.ffor_column_decode(O.FFORColumn(n_vecs * 1024, packed, bws, offs, bases))
    col = D.DeviceFFORColumn.from_host(packed, bws, offs, bases, n_vecs * 1024)
    return col, dec, bws, bases


def expected(dec, bws, bases, predicate, lo, hi):
    d = dec.astype(np.uint64).reshape(len(bws), 1024)
    if predicate == A.PRED_ALL:
        sel = np.ones(d.shape, bool)
    elif predicate == A.PRED_OFFSET_LOWER_HALF:
        off = (d - bases.astype(np.uint64)[:, None]) & np.uint64(0xFFFFFFFF)
        half = np.array([(1 << (int(b) - 1)) if b else 0 for b in bws], np.uint64)
        sel = off < half[:, None]
    else:
        sel = (d >= np.uint64(lo)) & (d < np.uint64(hi))
    return int(sel.sum()), int(d[sel].sum(dtype=np.uint64))


def test_noalloc_filter_sum_matches_allocating_entry_point():
    L = A.lib()
    rng = np.random.default_rng(2026)
    res_a = torch.zeros(2, dtype=torch.uint64, device="cuda")
    res_b = torch.zeros(2, dtype=torch.uint64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    # the first call's grid fills all 2048 partial slots; the second uses fewer
    for n_vecs, bw in ((40000, 13), (300, 17)):
        col, dec, bws, bases = column(rng, n_vecs, bw)
        st = col.ffor_struct()
        lo = int(dec[7])
        hi = min(lo + (1 << 12), 0xFFFFFFFF)
        for pred in (A.PRED_ALL, A.PRED_OFFSET_LOWER_HALF, A.PRED_RANGE):
            A.check(L.flsgpu_query_filter_sum_ffor_u32(C.byref(st), pred, lo, hi, res_a.data_ptr(),
                                                       0, stream))
            A.check(L.flsgpu_query_filter_sum_ffor_u32_noalloc(C.byref(st), pred, lo, hi,
                                                               res_b.data_ptr(), stream))
            torch.cuda.synchronize()
            a, b = res_a.cpu().numpy(), res_b.cpu().numpy()
            assert np.array_equal(a, b), (n_vecs, pred, a, b)
            c, s = expected(dec, bws, bases, pred, lo, hi)
            assert int(b[0]) == c and int(b[1]) == s, (n_vecs, pred)


def test_noalloc_rejects_bad_arguments():
    L = A.lib()
    ok = A.BPColumn(1024, 1, 0x1000, 0x2000, 0x3000)
    f = A.FFORColumn(ok, None)
    assert L.flsgpu_query_filter_sum_ffor_u32_noalloc(C.byref(f), 2, 0, 0, C.c_void_p(0x6000),
                                                      None) == A.FLSGPU_EINVAL
    f = A.FFORColumn(ok, 0x5000)
    assert L.flsgpu_query_filter_sum_ffor_u32_noalloc(C.byref(f), 2, 0, 0, None,
                                                      None) == A.FLSGPU_EINVAL
