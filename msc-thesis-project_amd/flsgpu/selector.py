"""Learned ALP parameter selector: training and evaluation driver.

The network and its training run in the library's HIP kernels
(include/flsgpu_selector.h): features, forward, backward, Adam, predict.
Labels come from the exhaustive per-vector (exponent, factor) search of the
library's encoder (flsgpu_alp_search_*), the same search the reference runs
on its samples (src/alp/encoder.hpp:145-197); a vector whose best estimate
reaches the ALP-RD threshold (encoder.hpp:208-211, 48 bits x 32 samples for
doubles) is labelled ALP-RD.

    python -m flsgpu.selector --train-log2 22 --out weights.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import time

import numpy as np
import torch

from . import _capi as A
from . import alp as AL

N_IN, N_H, N_OUT, N_PARAM = 32, 64, 40, 8872
RD_LIMIT_F64 = 48 * 32
_P = C.c_void_p


def _lib():
    L = A.lib()
    if not getattr(L, "_sel_sigs", False):
        for t in ("f32", "f64"):
            f = getattr(L, f"flsgpu_sel_features_{t}")
            f.argtypes = [_P, C.c_uint64, C.c_uint32, C.c_uint32, _P, _P]
            f.restype = C.c_int
        L.flsgpu_sel_forward.argtypes = [_P, _P, C.c_uint32, _P, _P, _P, _P]
        L.flsgpu_sel_backward.argtypes = [_P, _P, _P, _P, _P, _P, C.c_uint32, _P, _P, _P, _P, _P, _P]
        L.flsgpu_sel_adam.argtypes = [_P, _P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_uint32, _P]
        L.flsgpu_sel_predict.argtypes = [_P, _P, C.c_uint32, _P, _P, _P, _P]
        for t in ("f32", "f64"):
            getattr(L, f"flsgpu_sel_select_{t}").argtypes = [
                _P, _P, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P, _P, _P]
        for n in ("flsgpu_sel_forward", "flsgpu_sel_backward", "flsgpu_sel_adam",
                  "flsgpu_sel_predict", "flsgpu_sel_select_f32", "flsgpu_sel_select_f64"):
            getattr(L, n).restype = C.c_int
        L._sel_sigs = True
    return L


def _s():
    return torch.cuda.current_stream().cuda_stream


# ---------------------------------------------------------------------------
# Data: per-vector decimal precision 0..6, magnitude log-uniform in
# [1e-3, 1e6], 0..5% full-precision outliers; one vector in ten is uniform
# random doubles (ALP-RD territory).
# ---------------------------------------------------------------------------
def generate(n_vecs: int, rng, values_per_vec: int = 32) -> np.ndarray:
    d = rng.integers(0, 7, n_vecs)
    mag = 10.0 ** rng.uniform(-3.0, 6.0, n_vecs)
    u = rng.uniform(-1.0, 1.0, (n_vecs, values_per_vec))
    x = u * mag[:, None]
    scale = 10.0 ** d[:, None]
    x = np.round(x * scale) / scale
    p = rng.uniform(0.0, 0.05, n_vecs)
    outl = rng.random((n_vecs, values_per_vec)) < p[:, None]
    x[outl] = (rng.uniform(-1.0, 1.0, (n_vecs, values_per_vec)) * mag[:, None])[outl]
    rd = rng.random(n_vecs) < 0.1
    x[rd] = rng.uniform(0.0, 1.0, (int(rd.sum()), values_per_vec))
    return x


def label(samples: np.ndarray, n_threads: int = 0):
    """Exhaustive-search labels [n, 3] = (exponent, factor, scheme)."""
    e, f, est = AL.alp_search(samples, n_threads=n_threads)
    scheme = (est >= RD_LIMIT_F64).astype(np.int32)
    return np.stack([e.astype(np.int32), f.astype(np.int32), scheme], axis=1), est


def features(values: torch.Tensor, n_vecs: int, vector_stride: int, sample_stride: int) -> torch.Tensor:
    out = torch.empty((n_vecs, N_IN), dtype=torch.float32, device=values.device)
    fn = _lib().flsgpu_sel_features_f64 if values.dtype == torch.float64 else _lib().flsgpu_sel_features_f32
    A.check(fn(values.data_ptr(), n_vecs, vector_stride, sample_stride, out.data_ptr(), _s()))
    return out


def init_params(seed: int = 0, device="cuda") -> torch.Tensor:
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM, np.float32)
    o = 0
    for fan_in, fan_out in ((N_IN, N_H), (N_H, N_H), (N_H, N_OUT)):
        w = rng.normal(0.0, np.sqrt(2.0 / fan_in), (fan_in, fan_out)).astype(np.float32)
        p[o:o + w.size] = w.reshape(-1)
        o += w.size + fan_out  # bias stays 0
    return torch.from_numpy(p).to(device)


class Workspace:
    def __init__(self, batch: int, device="cuda"):
        z = lambda *s: torch.empty(s, dtype=torch.float32, device=device)  # noqa: E731
        self.h1, self.h2, self.y = z(batch, N_H), z(batch, N_H), z(batch, N_OUT)
        self.dy, self.dh1, self.dh2 = z(batch, N_OUT), z(batch, N_H), z(batch, N_H)
        self.grad, self.loss = z(N_PARAM), z(batch)


def forward_backward(params, x, labels, ws: Workspace):
    n = x.shape[0]
    L = _lib()
    A.check(L.flsgpu_sel_forward(params.data_ptr(), x.data_ptr(), n, ws.h1.data_ptr(),
                                 ws.h2.data_ptr(), ws.y.data_ptr(), _s()))
    A.check(L.flsgpu_sel_backward(params.data_ptr(), x.data_ptr(), ws.h1.data_ptr(),
                                  ws.h2.data_ptr(), ws.y.data_ptr(), labels.data_ptr(), n,
                                  ws.dy.data_ptr(), ws.dh1.data_ptr(), ws.dh2.data_ptr(),
                                  ws.grad.data_ptr(), ws.loss.data_ptr(), _s()))
    return ws.grad, ws.loss[:n]


def train(feats: torch.Tensor, labels: torch.Tensor, epochs: int = 3, batch: int = 4096,
          lr: float = 3e-3, seed: int = 0, log=None) -> torch.Tensor:
    dev = feats.device
    params = init_params(seed, dev)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    ws = Workspace(batch, dev)
    n = feats.shape[0]
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    step = 0
    L = _lib()
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(dev)
        tot = torch.zeros((), dtype=torch.float64, device=dev)
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            xb = feats[idx].contiguous()
            lb = labels[idx].contiguous()
            grad, loss = forward_backward(params, xb, lb, ws)
            tot += loss.double().sum()
            step += 1
            A.check(L.flsgpu_sel_adam(params.data_ptr(), grad.data_ptr(), m.data_ptr(),
                                      v.data_ptr(), lr, 0.9, 0.999, 1e-8, step, _s()))
        if log:
            log(f"[selector] epoch {ep}: mean loss {tot.item() / max(1, (n // batch) * batch):.4f}")
    return params


def predict(params: torch.Tensor, feats: torch.Tensor):
    n = feats.shape[0]
    e = torch.empty(n, dtype=torch.uint8, device=feats.device)
    f = torch.empty_like(e)
    s = torch.empty_like(e)
    A.check(_lib().flsgpu_sel_predict(params.data_ptr(), feats.data_ptr(), n, e.data_ptr(),
                                      f.data_ptr(), s.data_ptr(), _s()))
    return e, f, s


def select(params: torch.Tensor, feats: torch.Tensor, values: torch.Tensor, n_vecs: int,
           vector_stride: int, sample_stride: int, n_candidates: int = 4):
    """Learned proposal verified on the samples: the best of the network's
    n_candidates (exponent, factor) pairs under the search's size estimate.
    Returns exps, facs, scheme (uint8) and est_bits (int64) on the device."""
    dev = values.device
    e = torch.empty(n_vecs, dtype=torch.uint8, device=dev)
    f = torch.empty_like(e)
    s = torch.empty_like(e)
    est = torch.empty(n_vecs, dtype=torch.int64, device=dev)
    fn = _lib().flsgpu_sel_select_f64 if values.dtype == torch.float64 else _lib().flsgpu_sel_select_f32
    A.check(fn(params.data_ptr(), feats.data_ptr(), values.data_ptr(), n_vecs, vector_stride,
               sample_stride, n_candidates, e.data_ptr(), f.data_ptr(), s.data_ptr(),
               est.data_ptr(), _s()))
    return e, f, s, est


def evaluate(params: torch.Tensor, n_vecs: int = 4096, seed: int = 99, device="cuda",
             n_threads: int = 0, n_candidates: int = 4, reps: int = 10):
    """Compression ratio and selection time on a held-out column of n_vecs
    1024-value vectors: exhaustive search (CPU, the label source), the pure
    network proposal, the verified top-n_candidates proposal (GPU), and the
    reference's column-level sampling + top-k (CPU)."""
    rng = np.random.default_rng(seed)
    x = generate(n_vecs, rng, 1024).reshape(-1)
    samples = np.ascontiguousarray(x.reshape(n_vecs, 1024)[:, ::32])
    t0 = time.perf_counter()
    lab, est = label(samples, n_threads=1)
    t_search_1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    label(samples, n_threads=n_threads)
    t_search_mt = time.perf_counter() - t0
    xd = torch.from_numpy(x).to(device)

    def learned():
        fe = features(xd, n_vecs, 1024, 32)
        return predict(params, fe)

    def verified():
        fe = features(xd, n_vecs, 1024, 32)
        return select(params, fe, xd, n_vecs, 1024, 32, n_candidates)

    def timed(fn):
        out = fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return out, a.elapsed_time(b) / reps / 1e3

    (e, f, s), t_mlp = timed(learned)
    (ev, fv, sv, estv), t_ver = timed(verified)
    e, f, s = e.cpu().numpy(), f.cpu().numpy(), s.cpu().numpy()
    ev, fv, sv = ev.cpu().numpy(), fv.cpu().numpy(), sv.cpu().numpy()
    col_search = AL.alp_encode_with_params(x, lab[:, 0].astype(np.uint8), lab[:, 1].astype(np.uint8))
    col_mlp = AL.alp_encode_with_params(x, e, f)
    col_ver = AL.alp_encode_with_params(x, ev, fv)
    t0 = time.perf_counter()
    col_ref = AL.alp_encode(x, n_threads=1)
    t_ref = time.perf_counter() - t0
    raw = x.nbytes
    res = {
        "vectors": n_vecs,
        "n_candidates": n_candidates,
        "ratio_search": raw / col_search.compressed_bytes_alp,
        "ratio_learned": raw / col_mlp.compressed_bytes_alp,
        "ratio_verified": raw / col_ver.compressed_bytes_alp,
        "ratio_reference_topk": (raw / col_ref.compressed_bytes_alp) if col_ref else None,
        "exact_label_rate": float(((e == lab[:, 0]) & (f == lab[:, 1])).mean()),
        "exact_label_rate_verified": float(((ev == lab[:, 0]) & (fv == lab[:, 1])).mean()),
        "scheme_accuracy": float((s == lab[:, 2]).mean()),
        "scheme_accuracy_verified": float((sv == lab[:, 2]).mean()),
        "search_s_1thread": t_search_1,
        "search_s_threads": t_search_mt,
        "reference_encode_s_1thread": t_ref,
        "learned_s_gpu": t_mlp,
        "verified_s_gpu": t_ver,
    }
    res["ratio_loss_learned_vs_search"] = 1.0 - res["ratio_learned"] / res["ratio_search"]
    res["ratio_loss_verified_vs_search"] = 1.0 - res["ratio_verified"] / res["ratio_search"]
    res["speedup_verified_vs_search_1thread"] = t_search_1 / t_ver
    res["speedup_verified_vs_search_threads"] = t_search_mt / t_ver
    # lossless: decode the verified-proposal column on the GPU
    out = AL.decompress(AL.DeviceAlpColumn.from_host(col_ver, device))
    res["roundtrip_bit_exact"] = bool(torch.equal(out.view(torch.int64), xd.view(torch.int64)))
    return res


def train_default(train_log2: int = 22, epochs: int = 4, seed: int = 1, device="cuda", log=print,
                  batch: int = 4096):
    rng = np.random.default_rng(seed)
    n = 1 << train_log2
    chunks_f, chunks_l = [], []
    t0 = time.perf_counter()
    step = 1 << 20
    for c0 in range(0, n, step):
        m = min(step, n - c0)
        samples = generate(m, rng)
        lab, _ = label(samples)
        sd = torch.from_numpy(samples).to(device)
        chunks_f.append(features(sd, m, 32, 1))
        chunks_l.append(torch.from_numpy(lab).to(device))
    feats = torch.cat(chunks_f)
    labels = torch.cat(chunks_l)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    params = train(feats, labels, epochs=epochs, batch=batch, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    return params, {"train_vectors": n, "data_and_labels_s": t_data, "train_s": t_train,
                    "epochs": epochs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-log2", type=int, default=22)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    params, info = train_default(args.train_log2, args.epochs)
    res = evaluate(params)
    res.update(info)
    print(res)
    if args.out:
        np.savez(args.out, params=params.cpu().numpy())


if __name__ == "__main__":
    main()
