"""Region-scoring network kernels (needs a GPU): forward pass and gradients
against the FP64 CPU oracle on the same FP16-rounded weights; training
reduces the loss."""
import numpy as np
import pytest

import empi_amd as E
import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def test_forward_and_gradients_match_cpu_oracle():
    print("\n[gpu case] MLP forward/gradients vs FP64 oracle", flush=True)
    rng = np.random.default_rng(0)
    din = 80
    net = E.MLP(din, seed=3)
    p = net.params()
    # move biases away from zero so every term of the gradient is exercised
    p[256 * din:256 * din + 256] = rng.normal(0, 0.1, 256)
    net.set_params(p)
    X = rng.normal(0, 1, (200, din)).astype(np.float32)
    t = rng.normal(0, 1, 200).astype(np.float32)
    y = net.forward(X)
    y_ref, _, _ = mo.forward(p, X.astype(np.float64))
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    g, loss = net.gradients(X, t)
    g_ref, loss_ref = mo.gradients(p, X.astype(np.float64), t.astype(np.float64))
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    o = 0
    for name, n in [("W1", 256 * din), ("b1", 256), ("W2", 256 * 256), ("b2", 256), ("w3", 256), ("b3", 1)]:
        a, b = g[o:o + n], g_ref[o:o + n]
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-12)
        assert err < 1e-4, (name, err)
        o += n


def test_training_reduces_loss():
    print("\n[gpu case] MLP training", flush=True)
    rng = np.random.default_rng(1)
    din = 16
    X = rng.normal(0, 1, (20000, din)).astype(np.float32)
    t = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    net = E.MLP(din, seed=2)
    before = float(np.mean((net.forward(X[:2000]) - t[:2000]) ** 2))
    net.train(X, t, epochs=20, batch=1024, lr=2e-3)
    after = float(np.mean((net.forward(X[:2000]) - t[:2000]) ** 2))
    assert after < 0.1 * before, (before, after)


@pytest.mark.parametrize("din,rows", [(80, 200), (13, 45)])
def test_matrix_core_inference_matches_oracle(din, rows):
    """The proposal's inference pass (FP16 inputs and first hidden layer on
    v_mfma_f32_32x32x8f16) against the FP64 oracle with the same roundings,
    on the committed network and on an odd width / row count."""
    print(f"\n[gpu case] MLP matrix-core forward din={din} rows={rows}", flush=True)
    from empi_amd import proposal
    rng = np.random.default_rng(5 + din)
    if din == 80 and proposal.load() is not None:
        net = proposal.load()
        X = rng.normal(-8, 4, (rows, din)).astype(np.float32)
    else:
        net = E.MLP(din, seed=9)
        p = net.params()
        p[256 * din:256 * din + 256] = rng.normal(0, 0.1, 256)
        net.set_params(p)
        X = rng.normal(0, 1, (rows, din)).astype(np.float32)
    p = net.params()
    y = net.forward16(X)
    y_ref, _, _ = mo.forward(p, X.astype(np.float64), act16=True)
    scale = max(1.0, np.max(np.abs(y_ref)))
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * scale, np.max(np.abs(y - y_ref))
    # the FP32-activation pass differs by the roundings only
    y32 = net.forward(X)
    assert np.max(np.abs(y - y32)) <= 2e-2 * scale
