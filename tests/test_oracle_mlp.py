"""Pins of the learned-init encoder restatement (oracle.mlp_eval / oracle.mlp_adam).

The encoder has no counterpart in the reference (its lists start random), so these checks
fix it from mathematics: the analytic gradient against central finite differences of the
oracle's own FP64 loss, closed forms of the loss, Adam's constant-gradient closed form, and
the BF16 rounding variant staying close to plain FP64.
"""
import numpy as np
import pytest

import oracle


def _params(D, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(oracle.mlp_nparams(D))
    H, C = oracle.MLP_HID, oracle.MLP_CODE
    W1 = rng.uniform(-1, 1, (H, D + 1)) / np.sqrt(D) * scale
    W2 = rng.uniform(-1, 1, (C, H + 1)) / np.sqrt(H) * scale
    p[: H * (D + 1)] = W1.ravel()
    p[H * (D + 1):] = W2.ravel()
    return p


def _batch(B, P, D, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B * (1 + P), D)).astype(np.float32)
    # positives close to their anchors so both loss terms are active
    for b in range(B):
        for j in range(P):
            X[B + b * P + j] = X[b] + 0.3 * rng.standard_normal(D).astype(np.float32)
    return X


def test_gradient_matches_finite_differences():
    D, B, P = 32, 32, 2
    X = _batch(B, P, D, 1)
    mu = np.zeros(D, np.float32)
    prm = _params(D, 2, scale=3.0)
    margin = 4.0
    loss, g, _ = oracle.mlp_eval(X, B, P, mu, 1.0, prm, margin, round_bf16=False)
    assert loss > 0
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(oracle.MLP_HID * (D + 1), 40, replace=False),
                          oracle.MLP_HID * (D + 1) + rng.choice(oracle.MLP_CODE * 257, 40,
                                                                 replace=False)])
    h = 1e-6
    for i in idx:
        pp, pm = prm.copy(), prm.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.mlp_eval(X, B, P, mu, 1.0, pp, margin, False)[0]
              - oracle.mlp_eval(X, B, P, mu, 1.0, pm, margin, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_loss_closed_forms():
    D, B, P = 32, 32, 1
    X = _batch(B, P, D, 4)
    mu = np.zeros(D, np.float32)
    # all weights zero: every code equals b2, all distances vanish: loss = margin, grad of
    # W2 and b2 zero (dz = 0), grad of W1 zero
    prm = np.zeros(oracle.mlp_nparams(D))
    prm[oracle.MLP_HID * (D + 1) + oracle.MLP_HID::257] = 0.7  # b2
    loss, g, z = oracle.mlp_eval(X, B, P, mu, 1.0, prm, 2.5, False)
    assert loss == pytest.approx(2.5)
    assert np.all(z == 0.7)
    assert np.all(g == 0)
    # margin 0: the negative term vanishes; the loss is the mean squared positive distance
    prm = _params(D, 5)
    loss, _, z = oracle.mlp_eval(X, B, P, mu, 1.0, prm, 0.0, False)
    d = ((z[:B] - z[B:]) ** 2).sum(1)
    assert loss == pytest.approx(d.mean(), rel=1e-12)


def test_bf16_roundings_stay_close():
    D, B, P = 64, 64, 4
    X = _batch(B, P, D, 6)
    mu = X.mean(0)
    prm = _params(D, 7)
    l0, g0, z0 = oracle.mlp_eval(X, B, P, mu, 0.5, prm, 1.0, False)
    l1, g1, z1 = oracle.mlp_eval(X, B, P, mu, 0.5, prm, 1.0, True)
    assert abs(l1 - l0) < 0.02 * abs(l0)
    assert np.linalg.norm(g1 - g0) < 0.03 * np.linalg.norm(g0)
    assert np.abs(z1 - z0).max() < 0.02 * np.abs(z0).max()


def test_adam_constant_gradient_closed_form():
    # with a constant gradient the bias-corrected moments are g and g^2 exactly, so every
    # step moves by lr * g / (|g| + eps)
    n, lr = 1000, 1e-2
    rng = np.random.default_rng(8)
    g = rng.standard_normal(n)
    p0 = rng.standard_normal(n)
    p, m, v = p0.copy(), np.zeros(n), np.zeros(n)
    for t in range(1, 6):
        oracle.mlp_adam(t, lr, g, p, m, v)
    np.testing.assert_allclose(p, p0 - 5 * lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
