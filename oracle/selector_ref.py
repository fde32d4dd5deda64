"""CPU restatement of the learned ALP parameter selector (features, MLP
forward/backward in float64, Adam, predict).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
selector kernels, never as the product path.

The selector has no counterpart in the reference. Its labels come from the
exhaustive (exponent, factor) search of src/alp/encoder.hpp:145-197 (pinned
by the oracle's orc_alp_search_one_* and the reference build); the feature
round trip is the SAFE encode of encoder.hpp:76-84 followed by the decode of
decoder.hpp:121-124 at factor 0. What pins each function here:
  - roundtrip fractions: the oracle's own ALP encode/decode of single values
    (orc_alp_encode_value_*), checked in tests/test_selector.py;
  - MLP gradient: central finite differences in float64;
  - Adam: the closed form after one step (|update| = lr for every parameter
    with a non-zero gradient, since m/c1 = g and v/c2 = g^2).
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_H, N_OUT = 32, 64, 40
NE, NF, NS = 19, 19, 2
OW1 = 0
OB1 = OW1 + N_IN * N_H
OW2 = OB1 + N_H
OB2 = OW2 + N_H * N_H
OW3 = OB2 + N_H
OB3 = OW3 + N_H * N_OUT
N_PARAM = OB3 + N_OUT
assert N_PARAM == 8872

EXP10_F64 = np.array([10.0 ** k for k in range(19)], np.float64)
FRAC_F64 = np.array([float(f"1e-{k}") for k in range(21)], np.float64)
EXP10_F32 = np.array([10.0 ** k for k in range(11)], np.float32)
FRAC_F32 = np.array([float(f"1e-{k}") for k in range(11)], np.float64).astype(np.float32)
_LIM = 9223372036854774784.0


def roundtrip_mask(x: np.ndarray, k: int) -> np.ndarray:
    """Which values survive encode at (exponent k, factor 0) then decode."""
    if x.dtype == np.float64:
        with np.errstate(over="ignore", invalid="ignore"):
            t = x * EXP10_F64[k]
            bad = ~np.isfinite(t) | (t > _LIM) | (t < -_LIM) | ((t == 0.0) & np.signbit(t))
            t = np.where(bad, 0.0, t)
            t = (t + 6755399441055744.0) - 6755399441055744.0
            e = t.astype(np.int64)
            back = e.astype(np.float64) * FRAC_F64[k]
        return ~bad & (back == x)
    assert x.dtype == np.float32
    if k > 10:
        return np.zeros(x.shape, bool)
    with np.errstate(over="ignore", invalid="ignore"):
        t = x * EXP10_F32[k]
        td = t.astype(np.float64)
        bad = ~np.isfinite(t) | (td > _LIM) | (td < -_LIM) | ((t == 0.0) & np.signbit(t))
        t = np.where(bad, np.float32(0), t)
        t = (t + np.float32(12582912.0)) - np.float32(12582912.0)
        bad |= ~((t > np.float32(-2147483648.0)) & (t < np.float32(2147483648.0)))
        t = np.where(bad, np.float32(0), t)
        e = t.astype(np.int32)
        back = e.astype(np.float32) * FRAC_F32[k]
    return ~bad & (back == x)


def features(samples: np.ndarray) -> np.ndarray:
    """[n, 32] samples -> [n, 32] float32 features (definition of k_features)."""
    n = samples.shape[0]
    f = np.zeros((n, N_IN), np.float64)
    ok = np.stack([roundtrip_mask(samples, k).sum(axis=1) for k in range(19)], axis=1)
    f[:, :19] = ok / 32.0
    for r in range(n):
        s = samples[r]
        first_full = next((k for k in range(19) if ok[r, k] == 32), 19)
        mx, mn, sl, lo, hi, sm, sq = 0.0, 1e308, 0.0, 1e308, -1e308, 0.0, 0.0
        zeros = neg = special = nfin = same = distinct = 0
        for i in range(32):
            x = float(s[i])
            if not math.isfinite(x) or (x == 0.0 and math.copysign(1.0, x) < 0):
                special += 1
                continue
            nfin += 1
            a = abs(x)
            zeros += a == 0.0
            neg += x < 0.0
            if a > 0.0:
                mx = max(mx, a)
                mn = min(mn, a)
                sl += math.log10(a)
            lo = min(lo, x)
            hi = max(hi, x)
            sm += x
            sq += x * x
            same += s[i] == s[0]
            distinct += not any(s[j] == s[i] for j in range(i))
        nz = nfin - zeros
        f[r, 19] = math.log10(mx) / 10.0 if mx > 0.0 else -1.0
        f[r, 20] = math.log10(mn) / 10.0 if nz > 0 else -1.0
        f[r, 21] = sl / nz / 10.0 if nz > 0 else -1.0
        f[r, 22] = zeros / 32.0
        f[r, 23] = neg / 32.0
        f[r, 24] = special / 32.0
        if nfin > 1:
            d = hi - lo + 1.0
            f[r, 25] = min(math.log2(d) if math.isfinite(d) else math.inf, 1024.0) / 64.0
        k90 = next((k for k in range(19) if np.float32(f[r, k]) >= np.float32(0.9)), 19)
        f[r, 28] = k90 / 19.0
        f[r, 29] = first_full / 19.0
        f[r, 30] = 1.0
        var = max(sq / nfin - (sm / nfin) * (sm / nfin), 0.0) if nfin > 0 else 0.0
        if math.isnan(var):
            var = 0.0
        if var > 0.0:
            sd = math.sqrt(var) + 1e-300
            f[r, 31] = min(math.log10(sd) if math.isfinite(sd) else math.inf, 160.0) / 10.0
        else:
            f[r, 31] = -1.0
        f[r, 26] = same / 32.0
        f[r, 27] = distinct / 32.0
    return f.astype(np.float32)


def unpack(p: np.ndarray):
    p = p.astype(np.float64)
    return (p[OW1:OB1].reshape(N_IN, N_H), p[OB1:OW2], p[OW2:OB2].reshape(N_H, N_H), p[OB2:OW3],
            p[OW3:OB3].reshape(N_H, N_OUT), p[OB3:N_PARAM])


def forward(p: np.ndarray, x: np.ndarray):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = x.astype(np.float64)
    h1 = np.maximum(x @ w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2 + b2, 0.0)
    return h1, h2, h2 @ w3 + b3


_HEADS = ((0, NE), (NE, NF), (NE + NF, NS))


def loss_and_dy(y: np.ndarray, labels: np.ndarray):
    """Per-row summed cross-entropy of the three heads; dY of the mean loss."""
    n = y.shape[0]
    loss = np.zeros(n)
    dy = np.zeros_like(y)
    for h, (o, m) in enumerate(_HEADS):
        z = y[:, o:o + m]
        z = z - z.max(axis=1, keepdims=True)
        lse = np.log(np.exp(z).sum(axis=1))
        lab = labels[:, h]
        loss += lse - z[np.arange(n), lab]
        pr = np.exp(z - lse[:, None])
        pr[np.arange(n), lab] -= 1.0
        dy[:, o:o + m] = pr / n
    return loss, dy


def mean_loss(p, x, labels):
    return loss_and_dy(forward(p, x)[2], labels)[0].mean()


def backward(p: np.ndarray, x: np.ndarray, labels: np.ndarray):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = x.astype(np.float64)
    h1, h2, y = forward(p, x)
    loss, dy = loss_and_dy(y, labels)
    dh2 = (dy @ w3.T) * (h2 > 0)
    dh1 = (dh2 @ w2.T) * (h1 > 0)
    g = np.concatenate([(x.T @ dh1).ravel(), dh1.sum(0), (h1.T @ dh2).ravel(), dh2.sum(0),
                        (h2.T @ dy).ravel(), dy.sum(0)])
    return loss, g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1 = 1 - b1 ** step
    c2 = 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def predict_from_logits(y: np.ndarray):
    be = np.argmax(y[:, :NE], axis=1)
    bf = np.array([np.argmax(y[r, NE:NE + be[r] + 1]) for r in range(y.shape[0])])
    return be, bf, (y[:, NE + NF + 1] > y[:, NE + NF]).astype(np.int64)


MAXE = {np.dtype(np.float64): 18, np.dtype(np.float32): 10}
MAXF = {np.dtype(np.float64): 18, np.dtype(np.float32): 9}
RD_THRESHOLD = {np.dtype(np.float64): 48 * 32, np.dtype(np.float32): 22 * 32}


def candidates(y_row: np.ndarray, k: int, maxe: int, maxf: int):
    """The k (exponent, factor) pairs, factor <= min(exponent, maxf), with the
    largest y[e] + y[19 + f]; ties go to the pair met first in the order
    exponent descending, factor descending."""
    combos = [(e, f) for e in range(maxe, -1, -1) for f in range(min(e, maxf), -1, -1)]
    score = [float(np.float32(y_row[e]) + np.float32(y_row[NE + f])) for e, f in combos]
    taken = set()
    out = []
    for _ in range(min(k, len(combos))):
        best = None
        for i, (e, f) in enumerate(combos):
            if i in taken:
                continue
            if best is None or score[i] > score[best]:
                best = i
        taken.add(best)
        out.append(combos[best])
    return out


def select(y: np.ndarray, samples: np.ndarray, k: int):
    """Learned proposal + verification: the best of the k candidates under
    the search's own estimate and tie rules (encoder.hpp:182-190), starting
    from (0, 0) at the worst-case size. Returns exps, facs, est, scheme."""
    import oracle_py as O
    dt = samples.dtype
    excsz = 64 if dt == np.float64 else 32
    worst = 32 * (excsz + 16) + 32 * excsz
    n = samples.shape[0]
    E = np.zeros(n, np.int64)
    F = np.zeros(n, np.int64)
    EST = np.zeros(n, np.int64)
    for r in range(n):
        fe, ff, fest = 0, 0, worst
        for e, f in candidates(y[r], k, MAXE[dt], MAXF[dt]):
            est = O.alp_estimate(samples[r], e, f)
            if est is None:
                continue
            if est < fest or (est == fest and fe < e) or (est == fest and fe == e and ff < f):
                fe, ff, fest = e, f, est
        E[r], F[r], EST[r] = fe, ff, fest
    return E, F, EST, (EST >= RD_THRESHOLD[dt]).astype(np.int64)
