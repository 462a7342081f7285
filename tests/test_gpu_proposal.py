"""Learned proposal on the device (needs a GPU), judged by the exact search.

* delta = 0: every iteration is certified exact, so the atoms must pass the
  same FP64-oracle parity as the plain method.
* COLLECT mode runs the plain method's atoms and records training rows; the
  certified bound (feature 64, log bound / residual energy) must dominate
  every row's new exact maximum (its target).
* delta = 1e-3: explained-energy deficit against the plain run <= 0.1%,
  with fewer map entries evaluated.
"""

import numpy as np
import pytest

import empi_amd as E
import pyoracle as po
from empi_amd import synth
from test_gpu_parity import _announce, _check_item, _oracle

pytestmark = pytest.mark.gpu

LENGTHS = [1024, 1500, 2048, 3000]


def _signals(seed0=4100):
    return [synth.pink_with_atoms(seed0 + i, n, 256.0, 6) for i, n in enumerate(LENGTHS)]


def _plan(iters=25):
    return E.Plan([E.ItemSpec(1, n, 256.0, E.SMP, iters, 99.0, 0.01) for n in LENGTHS])


def _flat(xs):
    return np.concatenate([np.asarray(x, np.float32).ravel() for x in xs])


def _explained(plan, n):
    out = []
    for i in range(n):
        e = plan.result(i).energies[0]
        e = e[~np.isnan(e)]
        out.append(1.0 - e[-1] / e[0])
    return np.array(out)


def test_proposal_delta0_matches_oracle():
    _announce("proposal delta=0 (no network), uneven batch of 4, 25 iterations")
    xs = _signals()
    plan = _plan()
    plan.set_proposal(E.PROPOSAL_ON, None, top_k=8, delta=0.0)
    plan.run(_flat(xs))
    stats = plan.proposal_stats()
    print("stats", stats, flush=True)
    for i, x in enumerate(xs):
        res = plan.result(i)
        assert res.status == 0
        ora = _oracle(x[None, :], 256.0, 0.01, po.SMP, 25, 99.0)
        _check_item(res, ora, E.SMP, f"proposal item {i}")
    assert stats["certified_iterations"] > 0


def test_collect_rows_and_bound():
    _announce("proposal COLLECT mode: rows, targets below the certified bound")
    xs = _signals()
    plan = _plan()
    plan.set_proposal(E.PROPOSAL_COLLECT, None)
    plan.set_training_capacity(200000)
    plan.run(_flat(xs))
    X, t = plan.training_rows()
    print("rows", X.shape, flush=True)
    assert X.shape[0] > 100 and X.shape[1] == E.feature_dim()
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(t))
    # rows of regions whose new maximum is zero carry log(1e-30)
    live = t > -60
    assert np.all(t[live] <= X[live, 64] + 1e-3), float(np.max(t[live] - X[live, 64]))
    for i, x in enumerate(xs):
        res = plan.result(i)
        ora = _oracle(x[None, :], 256.0, 0.01, po.SMP, 25, 99.0)
        _check_item(res, ora, E.SMP, f"collect item {i}")
    # reading resets the store
    X2, _ = plan.training_rows()
    assert X2.shape[0] == 0


def test_proposal_delta_deficit_and_work():
    _announce("proposal delta=1e-3 vs plain: explained-energy deficit and entries")
    xs = _signals(4200)
    plain = _plan(40)
    plain.run(_flat(xs))
    e_plain = _explained(plain, len(xs))
    w_plain = plain.entries_evaluated()
    prop = _plan(40)
    prop.set_proposal(E.PROPOSAL_ON, None, top_k=8, delta=1e-3)
    prop.run(_flat(xs))
    e_prop = _explained(prop, len(xs))
    w_prop = prop.entries_evaluated()
    print("explained", e_plain, e_prop, "entries", w_plain, w_prop, prop.proposal_stats(), flush=True)
    deficit = (e_plain - e_prop) / e_plain
    assert np.all(deficit <= 1e-3), deficit
    assert w_prop < w_plain


def test_trained_network_gain():
    """The committed network (trained on the device on plain-method atoms of
    uneven batches with seeds 5000+i and 6000+i) against the method without
    it, on 24 config-4 signals (seeds 4000+i): ranking the dirty regions by
    the network, by their stale maximum, or not at all (top_k 0)."""
    _announce("trained network vs stale ranking vs none, 24 uneven signals, 100 iterations")
    from empi_amd import proposal
    net = proposal.load()
    assert net is not None, "empi_amd/proposal_net.npy missing"
    xs = synth.uneven_batch(items=24)
    specs = [E.ItemSpec(1, len(x), 256.0, E.SMP, 100, 99.0, 0.01) for x in xs]
    flat = _flat(xs)
    plain = E.Plan(specs)
    plain.run(flat)
    e_plain = _explained(plain, len(xs))
    out = {"plain": plain.entries_evaluated()}
    for name, k, n in [("none", 0, None), ("stale", 8, None), ("network", 8, net)]:
        plan = E.Plan(specs)
        plan.set_proposal(E.PROPOSAL_ON, n, top_k=k, delta=1e-4)
        plan.run(flat)
        deficit = (e_plain - _explained(plan, len(xs))) / e_plain
        out[name] = plan.entries_evaluated()
        print(name, f"entries {out[name]:.4g}", "stats", plan.proposal_stats(),
              "deficit max", float(deficit.max()), flush=True)
        assert float(deficit.max()) <= 1e-3, (name, deficit)
        plan.close()
    print("entries relative to plain:", {k: round(v / out["plain"], 4) for k, v in out.items()}, flush=True)
    assert out["network"] < out["none"]
    assert out["network"] < out["plain"]
    # the stated margin over ranking by stale maximum: at least 0.5% fewer
    # entries (2.3% measured: 0.540 against 0.553 of plain's entries)
    assert out["network"] <= 0.995 * out["stale"], (out["network"], out["stale"])
