"""The FP64 network oracle that checks the device kernels, pinned by central
finite differences of its own exact (unrounded) loss, and by a planted
mistake: a gradient without the second ReLU mask fails the same check."""
import numpy as np

import mlp_oracle as mo

H = mo.H


def _loss(p, X, t):
    y, _, _ = mo.forward(p, X, rounded=False)
    return float(np.mean((y - t) ** 2))


def _setup(seed=11, din=6, R=5):
    rng = np.random.default_rng(seed)
    n = H * din + H + H * H + H + H + 1
    p = rng.standard_normal(n) * 0.2
    X = rng.standard_normal((R, din))
    t = rng.standard_normal(R)
    return p, X, t, rng


def test_gradients_match_finite_differences():
    p, X, t, rng = _setup()
    g, loss = mo.gradients(p, X, t, rounded=False)
    assert abs(loss - _loss(p, X, t)) < 1e-12
    idx = rng.choice(p.size, 60, replace=False)
    idx = np.concatenate([idx, [p.size - 1, H * X.shape[1]]])   # b3 and b1[0]
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (_loss(p + e, X, t) - _loss(p - e, X, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_dropped_relu_mask_is_caught():
    p, X, t, rng = _setup()
    din = X.shape[1]
    W1, b1, W2, b2, w3, b3 = mo.unpack(p, din)
    y, h1, h2 = mo.forward(p, X, rounded=False)
    dy = 2.0 * (y - t) / len(t)
    dh2 = np.outer(dy, w3)                    # mistake: no (h2 > 0) mask
    dW2 = dh2.T @ h1
    g, _ = mo.gradients(p, X, t, rounded=False)
    o = H * din + H
    assert np.max(np.abs(dW2.ravel() - g[o:o + H * H])) > 1e-3
