"""Host side of the learned proposal (no device): folding the training
standardisation into the network is exact up to the FP16 weights the device
uses, checked with the FP64 network oracle."""
import numpy as np

import mlp_oracle as mo
from empi_amd import proposal


def test_fold_standardisation_matches_standardised_network():
    rng = np.random.default_rng(3)
    din = 80
    _, n = proposal._layout(din)
    p = (rng.standard_normal(n) * 0.05).astype(np.float32)
    mu = rng.uniform(-30, 5, din)
    sd = rng.uniform(0.05, 8, din)
    t_mu, t_sd = -7.5, 2.5
    X = mu + sd * rng.standard_normal((64, din))
    y_std, _, _ = mo.forward(p, (X - mu) / sd)
    folded = proposal.fold_standardisation(p, mu, sd, t_mu, t_sd)
    y_raw, _, _ = mo.forward(folded, X)
    want = t_mu + t_sd * y_std
    assert np.max(np.abs(y_raw - want)) <= 2e-3 * t_sd * max(1.0, np.max(np.abs(y_std)))
    # a dropped mean (planted on the CPU) is far outside that tolerance
    bad = proposal.fold_standardisation(p, np.zeros(din), sd, t_mu, t_sd)
    y_bad, _, _ = mo.forward(bad, X)
    assert np.max(np.abs(y_bad - want)) > 10 * np.max(np.abs(y_raw - want))


def test_layout_matches_oracle():
    din = 80
    o, n = proposal._layout(din)
    p = np.arange(n, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = mo.unpack(p, din)
    assert W1[0, 0] == o["W1"] and b1[0] == o["b1"] and W2[0, 0] == o["W2"]
    assert b2[0] == o["b2"] and w3[0] == o["w3"] and b3 == o["b3"] == n - 1
