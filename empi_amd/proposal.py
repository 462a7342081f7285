"""Learned proposal: training-row collection, device training of the
region-scoring network, weight files.

The network (empi_mlp_*, 80 -> 256 -> 256 -> 1) predicts, for a dirty region
(one scale x one 256-sample block), the log of its new exact maximum product
over the residual energy.  Rows come from COLLECT runs of the plain method
(every dirty region evaluated, so the atoms are the plain method's own).
Features and targets are standardised for training; the standardisation is
folded into the first and last layers afterwards, so inference reads the raw
rows the engine writes.
"""
from __future__ import annotations

import os

import numpy as np

from . import MLP, PROPOSAL_COLLECT, Plan, feature_dim

HIDDEN = 256
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "proposal_net.npy")
CLIP = -40.0          # log-energy floor (empty regions carry log 1e-30)


def collect(specs, flat_signal, capacity=2_000_000, stride=64, device_ptr=None):
    """Runs the plain method over the batch in COLLECT mode; returns (X, t)."""
    plan = Plan(specs)
    plan.set_proposal(PROPOSAL_COLLECT, None)
    plan.set_training_capacity(capacity, stride)
    if device_ptr is not None:
        plan.run(device_ptr=device_ptr)
    else:
        plan.run(flat_signal)
    X, t = plan.training_rows()
    atoms = sum(int(plan.result(i).counts.sum()) for i in range(len(specs)))
    plan.close()
    return X, t, atoms


def _layout(din):
    n1 = HIDDEN * din
    n2 = HIDDEN * HIDDEN
    o = {"W1": 0, "b1": n1, "W2": n1 + HIDDEN, "b2": n1 + HIDDEN + n2,
         "w3": n1 + 2 * HIDDEN + n2, "b3": n1 + 3 * HIDDEN + n2}
    return o, n1 + 3 * HIDDEN + n2 + 1


def fold_standardisation(params, mu, sd, t_mu, t_sd):
    """Parameters of a network trained on (X - mu) / sd -> (t - t_mu) / t_sd
    rewritten to read raw X and return raw t."""
    din = mu.size
    o, n = _layout(din)
    p = np.array(params, dtype=np.float64)
    W1 = p[o["W1"]:o["b1"]].reshape(HIDDEN, din)
    b1 = p[o["b1"]:o["W2"]]
    # the device rounds W1 to FP16: fold the mean through the rounded weights
    # so that only the standardised term carries the rounding
    W1n = (W1 / sd[None, :]).astype(np.float32).astype(np.float16).astype(np.float64)
    b1n = b1 - W1n @ mu
    p[o["W1"]:o["b1"]] = W1n.ravel()
    p[o["b1"]:o["W2"]] = b1n
    p[o["w3"]:o["b3"]] *= t_sd
    p[o["b3"]] = p[o["b3"]] * t_sd + t_mu
    return p.astype(np.float32)


def fit(X, t, epochs=4, batch=4096, lr=1e-3, seed=0):
    """Trains a fresh network on the device; returns (MLP, final loss on the
    standardised targets)."""
    X = np.maximum(np.asarray(X, np.float32), CLIP)
    t = np.maximum(np.asarray(t, np.float32), CLIP)
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    sd[sd < 1e-6] = 1.0
    t_mu, t_sd = float(t.mean()), float(t.std() or 1.0)
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(t))
    Xs = ((X[order] - mu) / sd).astype(np.float32)
    ts = ((t[order] - t_mu) / t_sd).astype(np.float32)
    net = MLP(feature_dim(), seed=seed)
    loss = net.train(Xs, ts, epochs=epochs, batch=batch, lr=lr, seed=seed)
    net.set_params(fold_standardisation(net.params(), mu, sd, t_mu, t_sd))
    return net, loss


def save(net, path=WEIGHTS):
    np.save(path, net.params())


def load(path=WEIGHTS):
    """The network stored at path, or None when there is no file."""
    if not os.path.exists(path):
        return None
    p = np.load(path, allow_pickle=False)
    net = MLP(feature_dim())
    net.set_params(p)
    return net


def training_set(batches=(5000, 6000), items=512, iterations=100, stride=64, capacity=4_000_000, log=print):
    """Rows of COLLECT runs over uneven batches (lengths U{1024..16384} at
    256 Hz, pink noise + Poisson(6) atoms, energyError 0.01, 99% / at most
    `iterations`), seeds b+i for each b in batches (disjoint from the
    benchmark's 4000+i)."""
    from . import ItemSpec, SMP
    from .synth import uneven_batch
    Xs, ts, total = [], [], 0
    for b in batches:
        xs = uneven_batch(items=items, seed0=b)
        specs = [ItemSpec(1, len(x), 256.0, SMP, iterations, 99.0, 0.01) for x in xs]
        flat = np.concatenate([x.astype(np.float32) for x in xs])
        X, t, atoms = collect(specs, flat, capacity=capacity, stride=stride)
        log(f"batch {b}: {atoms} atoms, {len(t)} rows")
        Xs.append(X)
        ts.append(t)
        total += atoms
    return np.concatenate(Xs), np.concatenate(ts), total


def _explained(plan, n):
    """(explained fraction per signal, explained energy per signal)."""
    frac, energy = [], []
    for i in range(n):
        e = plan.result(i).energies[0]
        e = e[~np.isnan(e)]
        frac.append(1.0 - e[-1] / e[0])
        energy.append(e[0] - e[-1])
    return np.array(frac), np.array(energy)


def sweep(deltas, items=512, top_k=8, log=print):
    """Atoms/s and explained-energy deficit of the proposal against the plain
    method over a range of delta (the per-iteration tolerance: every chosen
    atom's product is at least (1 - delta) times the exact maximum wherever
    the bounds hold), on the uneven benchmark batch (512 signals of
    1024..16384 samples at 256 Hz, seeds 4000+i).  Each run is timed after an
    untimed one.  Deficits per signal (max, mean) and of the batch (explained
    energy summed over signals)."""
    import time
    import torch
    from . import ItemSpec, PROPOSAL_ON, SMP
    from .synth import uneven_batch
    xs = uneven_batch(items=items)
    specs = [ItemSpec(1, len(x), 256.0, SMP, 100, 99.0, 0.01) for x in xs]
    dev = torch.from_numpy(np.concatenate([x.astype(np.float32) for x in xs])).to("cuda:0")
    net = load()
    tuning = {}   # plain, proposal: each method's own autotuned table (the proposal runs low-bin kernels)

    def timed(delta):
        plan = Plan(specs)
        key = delta is not None
        if delta is not None:
            plan.set_proposal(PROPOSAL_ON, net, top_k=top_k, delta=delta)
        if key not in tuning:
            tuning[key] = plan.autotune(device_ptr=dev.data_ptr(), repeats=2)
        else:
            plan.set_tuning(tuning[key])
        plan.run(device_ptr=dev.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run(device_ptr=dev.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        atoms = sum(int(plan.result(i).counts.sum()) for i in range(len(specs)))
        frac, energy = _explained(plan, len(specs))
        entries = plan.entries_evaluated()
        plan.close()
        return atoms / dt, entries, frac, energy

    base_rate, base_entries, base_frac, base_energy = timed(None)
    log(f"plain: {base_rate:.0f} atoms/s")
    rows = []
    for delta in deltas:
        rate, entries, frac, energy = timed(delta)
        d = (base_frac - frac) / base_frac
        row = {"delta": delta, "atoms_per_s": rate, "speedup": rate / base_rate,
               "entries_ratio": base_entries / entries, "deficit_max": float(d.max()),
               "deficit_mean": float(d.mean()),
               "deficit_batch": float((base_energy.sum() - energy.sum()) / base_energy.sum())}
        log(row)
        rows.append(row)
    return base_rate, rows


def main():
    import argparse
    import time
    ap = argparse.ArgumentParser(description="train the proposal network on the device, or sweep delta")
    ap.add_argument("--items", type=int, default=512)
    ap.add_argument("--stride", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--sweep", default=None,
                    help="comma-separated deltas: measure speed and deficit against the plain method instead of training")
    a = ap.parse_args()
    if a.sweep:
        sweep([float(v) for v in a.sweep.split(",")], items=a.items, log=lambda m: print(m, flush=True))
        return
    t0 = time.perf_counter()
    X, t, atoms = training_set(items=a.items, stride=a.stride, log=lambda m: print(m, flush=True))
    print(f"collected {len(t)} rows from {atoms} atoms in {time.perf_counter() - t0:.1f} s", flush=True)
    t1 = time.perf_counter()
    net, loss = fit(X, t, epochs=a.epochs)
    print(f"trained {a.epochs} epochs in {time.perf_counter() - t1:.1f} s, standardised loss {loss:.4f}", flush=True)
    save(net, a.out)
    print(f"wrote {a.out}", flush=True)


if __name__ == "__main__":
    main()
