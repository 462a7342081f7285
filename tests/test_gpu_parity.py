"""HIP engine vs the FP64 oracle on the same seeded inputs (needs a GPU).

Each case prints its name before touching the GPU.  Sizes are chosen so the
oracle finishes in seconds.
"""
import sys

import numpy as np
import pytest

import empi_amd as E
import pyoracle as po
from empi_amd import synth
from parity import ENERGY_RTOL, assert_parity, compare_rows, energy_rel, oracle_following

pytestmark = pytest.mark.gpu


def _announce(name):
    print(f"\n[gpu case] {name}", flush=True)
    sys.stdout.flush()


class _OracleCase:
    """An FP64-oracle run of one item, made when the engine's result is known
    so that the oracle can follow it through near-ties."""

    def __init__(self, x, fs, eps, variant, iters, percent, kw):
        self.x, self.fs, self.eps, self.variant = np.asarray(x, np.float64), fs, eps, variant
        self.iters, self.percent, self.kw = iters, percent, kw

    def run(self, res):
        d = po.Dictionary(self.fs, self.x.shape[-1], self.eps, **self.kw)
        return oracle_following(d, self.x, self.variant, self.iters, self.percent, res.index, res.counts)


def _oracle(x, fs, eps, variant, iters, percent, **kw):
    return _OracleCase(x, fs, eps, variant, iters, percent, kw)


def _check_item(res, case, variant, where, min_compared=1):
    """res: E.ItemResult; case: _oracle(...) of the same signal.  Every
    iteration is compared (parity.py)."""
    ora = case.run(res)
    C = res.params.shape[0]
    rows = C if variant == E.SMP else 1
    report = []
    for c in range(C):
        r = c if variant == E.SMP else 0
        err = compare_rows(res.params[c], res.index[c], res.counts[c], ora.params[c], ora.index[c],
                           ora.counts[c], ora.followed[r])
        assert_parity(err, f"{where} channel {c}", min_compared)
        report.append(err)
    for r in range(rows):
        er = energy_rel(res.energies[r], ora.energies[r], int(ora.counts[r]))
        assert er <= ENERGY_RTOL, (where, r, er)
    print(f"{where}: compared {[e['n'] for e in report]} atoms, near-ties followed "
          f"{[e['followed'] for e in report]}", flush=True)
    return report


def _run_plan(specs, signals):
    plan = E.Plan(specs)
    plan.run(np.concatenate([np.asarray(s, dtype=np.float32).ravel() for s in signals]))
    return plan


def test_smp_single_channel_1280():
    _announce("SMP 1280 samples @128 Hz, eps 0.01, 50 iterations, 99%")
    x = synth.single_channel()
    spec = E.ItemSpec(1, 1280, 128.0, E.SMP, 50, 99.0, 0.01)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(x[None, :], 128.0, 0.01, po.SMP, 50, 99.0)
    rep = _check_item(res, ora, E.SMP, "config1")
    print("errors", rep)
    assert plan.atom_count(0) == 5008073


def test_smp_batch_pink():
    _announce("SMP batch of 3 channels, 1024 samples @256 Hz, 30 iterations")
    xs = np.stack([synth.pink_with_atoms(2000 + i, 1024, 256.0, 8) for i in range(3)])
    spec = E.ItemSpec(3, 1024, 256.0, E.SMP, 30, 99.0, 0.01)
    plan = _run_plan([spec], [xs])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(xs, 256.0, 0.01, po.SMP, 30, 99.0)
    _check_item(res, ora, E.SMP, "smp batch")


@pytest.mark.parametrize("variant", [E.MMP1, E.MMP2, E.MMP3])
def test_multichannel_variants(variant):
    _announce(f"variant {variant}: 4 channels, 1280 samples @128 Hz, 20 iterations")
    x = synth.shared_multichannel(C=4, N=1280, fs=128.0, atoms=4, noise_sigma=0.1, seed=3100 + variant)
    spec = E.ItemSpec(4, 1280, 128.0, variant, 20, 99.5, 0.01)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(x, 128.0, 0.01, variant, 20, 99.5)
    _check_item(res, ora, variant, f"variant {variant}")


def test_uneven_batch():
    _announce("uneven batch: 4 signals of 1024..3000 samples @256 Hz")
    lengths = [1024, 1500, 2048, 3000]
    xs = [synth.pink_with_atoms(4100 + i, n, 256.0, 6) for i, n in enumerate(lengths)]
    specs = [E.ItemSpec(1, n, 256.0, E.SMP, 25, 99.0, 0.01) for n in lengths]
    plan = _run_plan(specs, xs)
    for i, x in enumerate(xs):
        res = plan.result(i)
        assert res.status == 0
        ora = _oracle(x[None, :], 256.0, 0.01, po.SMP, 25, 99.0)
        _check_item(res, ora, E.SMP, f"uneven item {i}")


def test_workspace_api_reference_loop():
    """The reference's Decomposition::compute loop (classes.cpp:26-73) run on
    the host over the per-call Workspace ABI."""
    _announce("workspace ABI: reference loop on the host, config-1 signal, 12 iterations")
    x = synth.single_channel().astype(np.float64)
    ws = E.Workspace(128.0, 1, 1280, 0.01)
    assert ws.atom_count() == 5008073
    ws.compute(x[None, :])
    residue = x.copy()
    total = float(np.sum(x * x) / 128.0)
    atoms = []
    for it in range(1, 13):
        best = ws.find_best()
        atoms.append(best[0])
        if it == 12:
            break
        ws.subtract(best[0], 0, residue)
        if float(np.sum(residue * residue) / 128.0) / total <= 0.01:
            break
    params = np.array([list(a.params) for a in atoms])
    index = np.array([(a.scale, a.freq, a.time) for a in atoms])
    d = po.Dictionary(128.0, 1280, 0.01)
    ora = oracle_following(d, x[None, :], po.SMP, 12, 99.0, index[None], [len(atoms)])
    err = compare_rows(params, index, len(atoms), ora.params[0], ora.index[0], ora.counts[0], ora.followed[0])
    assert_parity(err, "workspace", min_compared=len(atoms))


def test_signal_is_empty_status():
    _announce("all-zero signal -> signalIsEmpty")
    spec = E.ItemSpec(1, 1024, 256.0, E.SMP, 5, 99.0, 0.01)
    plan = _run_plan([spec], [np.zeros(1024, np.float32)])
    assert plan.result(0).status == 1
