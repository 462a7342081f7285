"""FP64 restatement of the region-scoring network (project-original, no
reference counterpart).  TEST INFRASTRUCTURE ONLY: checks the device forward
pass and gradients of empi_mlp_* on the same FP16-rounded weights.

Layout of the flat parameter block: W1 [256][din], b1 [256], W2 [256][256],
b2 [256], w3 [256], b3 [1].  Forward: relu(W1 x + b1) -> relu(W2 h1 + b2)
-> w3 . h2 + b3.  Loss: mean squared error.
"""
import numpy as np

H = 256


def unpack(p, din):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * din].reshape(H, din); o += H * din
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def half(a):
    return np.asarray(a, dtype=np.float32).astype(np.float16).astype(np.float64)


def forward(p, X, rounded=True, act16=False):
    """rounded: weights through FP16 and biases through FP32, as the device
    reads them; False gives the exact FP64 network (finite-difference checks).
    act16: inputs and first hidden layer also through FP16, as the
    matrix-core inference pass reads them."""
    din = X.shape[1]
    W1, b1, W2, b2, w3, b3 = unpack(p, din)
    if rounded:
        W1, W2, w3 = half(W1), half(W2), half(w3)
        b1 = b1.astype(np.float32).astype(np.float64)
        b2 = b2.astype(np.float32).astype(np.float64)
        b3 = np.float64(np.float32(b3))
    if act16:
        X = half(X)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    if act16:
        h1 = half(h1)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ w3 + b3, h1, h2


def gradients(p, X, t, rounded=True):
    """(flat gradient, loss) of mean((y - t)^2); with rounded weights the
    rounding is passed straight through (the device's training rule)."""
    R, din = X.shape
    W1, b1, W2, b2, w3, b3 = unpack(p, din)
    W2h, w3h = (half(W2), half(w3)) if rounded else (W2, w3)
    y, h1, h2 = forward(p, X, rounded)
    dy = 2.0 * (y - t) / R
    dw3 = h2.T @ dy
    db3 = dy.sum()
    dh2 = np.outer(dy, w3h) * (h2 > 0)
    dW2 = dh2.T @ h1
    db2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2h) * (h1 > 0)
    dW1 = dh1.T @ X
    db1 = dh1.sum(axis=0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return g, float(np.mean((y - t) ** 2))
