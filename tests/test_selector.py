"""CPU checks of the selector's oracle (oracle/selector_ref.py): the
round-trip features against the oracle's ALP vector encoder, the float64
gradient against central finite differences, Adam against its one-step
closed form, and the training labels against the oracle's exhaustive search."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle_py as O  # noqa: E402
import selector_ref as S  # noqa: E402


def _vectors(dtype, rng, n=4):
    out = []
    for d in range(n):
        x = np.round(rng.uniform(-1e4, 1e4, 1024), d).astype(dtype)
        x[rng.choice(1024, 20, replace=False)] = rng.uniform(-1, 1, 20).astype(dtype)
        out.append(x)
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_roundtrip_mask_matches_oracle_encoder(dtype):
    rng = np.random.default_rng(11)
    kmax = 19 if dtype == np.float64 else 11
    for x in _vectors(dtype, rng):
        for k in range(kmax):
            _, _, pos = O.alp_encode_vec(x, 0, k)
            want = np.ones(1024, bool)
            want[pos] = False
            got = S.roundtrip_mask(x, k)
            assert np.array_equal(got, want), (dtype, k)


def test_roundtrip_mask_special_values():
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e300, -1e300, 0.5, 2.0 ** 62, 2.0 ** 63])
    m = S.roundtrip_mask(x, 0)
    assert m.tolist() == [True, False, False, False, False, False, False, False, True, False]
    m2 = S.roundtrip_mask(x, 1)
    assert m2[7] and not m2[1]


def test_features_closed_forms():
    s = np.zeros((3, 32))
    s[0] = np.arange(32) * 0.25                      # two decimals suffice
    s[1] = 7.0                                       # constant
    s[2, :16] = -0.0
    s[2, 16:] = np.nan
    f = S.features(s)
    assert f[0, 0] == pytest.approx(8 / 32) and f[0, 2] == 1.0   # 0.25*k integral for k%4==0
    assert f[0, 29] == pytest.approx(2 / 19)                       # first exponent with all 32
    assert f[0, 22] == pytest.approx(1 / 32) and f[0, 27] == 1.0  # one zero, all distinct
    assert f[1, 26] == 1.0 and f[1, 27] == pytest.approx(1 / 32) and f[1, 31] == -1.0
    assert f[1, 19] == pytest.approx(np.log10(7) / 10, rel=1e-6)
    assert f[2, 24] == 1.0 and f[2, 19] == -1.0 and f[2, 25] == 0.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n = 12
    x = rng.normal(size=(n, S.N_IN))
    labels = np.stack([rng.integers(0, 19, n), rng.integers(0, 19, n), rng.integers(0, 2, n)], 1)
    p = rng.normal(0, 0.3, S.N_PARAM)
    _, g = S.backward(p, x, labels)
    idx = np.concatenate([rng.choice(np.arange(a, b), 12, replace=False) for a, b in
                          ((S.OW1, S.OB1), (S.OB1, S.OW2), (S.OW2, S.OB2), (S.OB2, S.OW3),
                           (S.OW3, S.OB3), (S.OB3, S.N_PARAM))])
    h = 1e-6
    for i in idx:
        pp = p.copy()
        pm = p.copy()
        pp[i] += h
        pm[i] -= h
        num = (S.mean_loss(pp, x, labels) - S.mean_loss(pm, x, labels)) / (2 * h)
        assert num == pytest.approx(g[i], rel=1e-5, abs=1e-8), i


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p = rng.normal(size=S.N_PARAM)
    g = rng.normal(size=S.N_PARAM)
    g[::7] = 0.0
    p1, m, v = S.adam(p, g, np.zeros_like(p), np.zeros_like(p), 1e-3, 0.9, 0.999, 1e-30, 1)
    d = p - p1
    nz = g != 0
    assert np.allclose(np.abs(d[nz]), 1e-3, rtol=1e-9)
    assert np.all(np.sign(d[nz]) == np.sign(g[nz])) and np.all(d[~nz] == 0)


def test_predict_factor_never_exceeds_exponent():
    rng = np.random.default_rng(5)
    y = rng.normal(size=(200, S.N_OUT))
    e, f, s = S.predict_from_logits(y)
    assert np.all(f <= e)
    assert np.array_equal(s, (y[:, 39] > y[:, 38]).astype(int))


def test_labels_match_oracle_search():
    """The training labels (product encoder's exhaustive search) equal the
    oracle's search on the same samples, scheme threshold included."""
    from flsgpu import selector as SEL
    rng = np.random.default_rng(6)
    samples = SEL.generate(300, rng)
    lab, est = SEL.label(samples)
    for r in range(samples.shape[0]):
        e, f, est_o = O.alp_search_one(samples[r])
        assert (lab[r, 0], lab[r, 1], est[r]) == (e, f, est_o), r
        assert lab[r, 2] == int(est_o >= 48 * 32)
    assert 0 < lab[:, 2].sum() < len(lab)
