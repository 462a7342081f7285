"""Benchmark workloads at their full sizes (needs a GPU), one plan each:

* the default SMP batch, 64 channels x 20 epochs x 5120 samples at 256 Hz,
  energyError 0.01, 100 iterations: every channel-epoch has status OK, 100
  atoms (99% is not reached on pink noise), positive non-increasing residual
  energies and atom indices inside the dictionary; three channel-epochs (the
  first, one in the middle, the last) match the FP64 oracle over their first
  30 iterations;
* MMP3, 32 channels x 10240 samples at 512 Hz, 200 iterations: the same
  properties, and one (scale, bin, time) per iteration across channels;
* the uneven batch, 512 signals of 1024..16384 samples: the same properties
  for every signal, the two shortest against the oracle.
"""
import numpy as np
import pytest

import empi_amd as E
import pyoracle as po
from empi_amd import synth
from parity import ENERGY_RTOL, assert_parity, compare_rows, compared_iterations, energy_rel
from test_gpu_parity import _announce

pytestmark = pytest.mark.gpu

CH, EPOCHS, N, FS, ITERS = 64, 20, 5120, 256.0, 100


def test_default_batch_full_size():
    _announce(f"full size: {CH} ch x {EPOCHS} epochs x {N} samples @{FS:g} Hz, {ITERS} iterations")
    x = synth.batched_channels(items=CH * EPOCHS, N=N, fs=FS)
    specs = [E.ItemSpec(CH, N, FS, E.SMP, ITERS, 99.0, 0.01) for _ in range(EPOCHS)]
    plan = E.Plan(specs)
    plan.run(x.ravel())
    d = po.Dictionary(FS, N, 0.01)
    L = len(d.scales)
    for e in range(EPOCHS):
        res = plan.result(e)
        assert res.status == 0, e
        assert np.all(res.counts == ITERS), (e, res.counts)
        en = res.energies[:, :ITERS]
        assert np.all(en > 0), e
        assert np.all(np.diff(en, axis=1) <= 1e-12 * en[:, :1]), e
        k, f, t = res.index[..., 0], res.index[..., 1], res.index[..., 2]
        assert np.all((k >= 0) & (k < L)), e
        F = np.array([sc.F for sc in d.scales])
        T = np.array([sc.T for sc in d.scales])
        assert np.all((f >= 0) & (f < F[k]) & (t >= 0) & (t < T[k])), e

    n_ora = 30
    for e, c in [(0, 0), (EPOCHS // 2, CH // 2), (EPOCHS - 1, CH - 1)]:
        res = plan.result(e)
        ora = po.decompose(d, x[e * CH + c][None, :].astype(np.float64), po.SMP, n_ora, 99.0, threads=0)
        n = compared_iterations(ora.gaps[0], int(ora.counts[0]))
        err = compare_rows(res.params[c], res.index[c], ora.params[0], ora.index[0], n)
        print(f"epoch {e} channel {c}: {n} compared, {err}", flush=True)
        assert_parity(err, f"epoch {e} channel {c}")
        assert energy_rel(res.energies[c], ora.energies[0], n) <= ENERGY_RTOL


def _check_properties(res, d, iters, rows):
    assert res.status == 0
    en = res.energies[:rows]
    n = int(res.counts.max())
    assert 0 < n <= iters
    e = en[:, :n]
    ok = ~np.isnan(e)
    assert np.all(e[ok] > 0)
    for r in range(rows):
        v = e[r][ok[r]]
        assert np.all(np.diff(v) <= 1e-12 * v[0]), r
    k, f, t = res.index[..., 0], res.index[..., 1], res.index[..., 2]
    live = np.arange(res.index.shape[1])[None, :] < res.counts[:, None]
    F = np.array([sc.F for sc in d.scales])
    T = np.array([sc.T for sc in d.scales])
    kk = np.where(live, k, 0)
    assert np.all(~live | ((k >= 0) & (k < len(d.scales)) & (f >= 0) & (f < F[kk]) & (t >= 0) & (t < T[kk])))


def test_mmp3_full_size():
    """MMP3 at its full size: 32 channels x 10240 samples at 512 Hz, 12 shared
    atoms, energyError 0.005, 200 iterations (properties only: the oracle
    needs half a minute per iteration at this size)."""
    _announce("full size: MMP3 32 ch x 10240 samples @512 Hz, eps 0.005, 200 iterations")
    x = synth.shared_multichannel()
    spec = E.ItemSpec(32, 10240, 512.0, E.MMP3, 200, 99.0, 0.005)
    plan = E.Plan([spec])
    plan.run(x.ravel())
    res = plan.result(0)
    _check_properties(res, po.Dictionary(512.0, 10240, 0.005), 200, 1)
    # MMP3 shares scale, frequency and position across channels
    n = int(res.counts[0])
    assert np.all(res.counts == n)
    assert np.all(res.index[:, :n] == res.index[:1, :n])


def test_uneven_batch_full_size():
    """The uneven batch at its full size: 512 signals of 1024..16384 samples at
    256 Hz, 99% / at most 100 iterations, one plan.  Properties for every
    signal; the two shortest against the FP64 oracle over 30 iterations."""
    _announce("full size: uneven batch, 512 signals of 1024..16384 samples @256 Hz, 100 iterations")
    xs = synth.uneven_batch()
    specs = [E.ItemSpec(1, len(s), 256.0, E.SMP, 100, 99.0, 0.01) for s in xs]
    plan = E.Plan(specs)
    plan.run(np.concatenate(xs))
    dicts = {}
    for i, s in enumerate(xs):
        d = dicts.setdefault(len(s), po.Dictionary(256.0, len(s), 0.01))
        _check_properties(plan.result(i), d, 100, 1)
    for i in np.argsort([len(s) for s in xs])[:2]:
        res = plan.result(int(i))
        d = dicts[len(xs[i])]
        ora = po.decompose(d, xs[i][None, :].astype(np.float64), po.SMP, 30, 99.0, threads=0)
        n = compared_iterations(ora.gaps[0], int(ora.counts[0]))
        err = compare_rows(res.params[0], res.index[0], ora.params[0], ora.index[0], n)
        print(f"signal {i} ({len(xs[i])} samples): {n} compared, {err}", flush=True)
        assert_parity(err, f"uneven signal {i}")
        assert energy_rel(res.energies[0], ora.energies[0], n) <= ENERGY_RTOL
