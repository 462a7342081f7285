"""Benchmark workloads at their full sizes (needs a GPU), one plan each.

Every unit of every run is checked for: status OK, positive non-increasing
residual energies, atom indices inside the dictionary, and at every
iteration the energy-drop identity fs (E_i - E_{i+1}) = sum_c (2 modulus^2 -
||w||^2) (the chosen map value gives <r, w> = modulus^2 exactly), which ties
each atom's FP64 parameters to the residual the engine actually subtracted.
Sampled units are decomposed by the FP64 oracle following the engine through
near-ties, and compared at every iteration (parity.py):

* the default SMP batch, 64 channels x 20 epochs x 5120 samples at 256 Hz,
  energyError 0.01, 100 iterations: three channel-epochs over 30 iterations
  and the last one over all 100;
* MMP3, 32 channels x 10240 samples at 512 Hz, energyError 0.005, 200
  iterations; the oracle on the same dictionary with the first 4 channels
  over 3 iterations (MMP3 at full channel count needs 80 GB of FP64 maps);
* the uneven batch, 512 signals of 1024..16384 samples: the two shortest
  over all their iterations and the median one over 30.
"""
import numpy as np
import pytest

import empi_amd as E
import pyoracle as po
from empi_amd import synth
from parity import ENERGY_RTOL, assert_parity, compare_rows, energy_drop_identity, energy_rel, oracle_following
from test_gpu_parity import _announce

pytestmark = pytest.mark.gpu

CH, EPOCHS, N, FS, ITERS = 64, 20, 5120, 256.0, 100
IDENTITY_RTOL = 1e-7


def _check_properties(res, d, iters, rows, where):
    """Status, counts, energies and index ranges of one item, and the
    energy-drop identity of every row."""
    assert res.status == 0, where
    n = int(res.counts.max())
    assert 0 < n <= iters, where
    e = res.energies[:rows, :n]
    ok = ~np.isnan(e)
    assert np.all(e[ok] > 0), where
    for r in range(rows):
        v = e[r][ok[r]]
        assert np.all(np.diff(v) <= 1e-12 * v[0]), (where, r)
    k, f, t = res.index[..., 0], res.index[..., 1], res.index[..., 2]
    live = np.arange(res.index.shape[1])[None, :] < res.counts[:, None]
    F = np.array([sc.F for sc in d.scales])
    T = np.array([sc.T for sc in d.scales])
    kk = np.where(live, k, 0)
    assert np.all(~live | ((k >= 0) & (k < len(d.scales)) & (f >= 0) & (f < F[kk]) & (t >= 0) & (t < T[kk]))), where
    worst = 0.0
    for r in range(rows):
        params = res.params[r] if rows > 1 else res.params
        cnt = int(res.counts[r])
        worst = max(worst, energy_drop_identity(d.fs, d.N, params[..., :cnt, :], res.energies[r], cnt))
    assert worst <= IDENTITY_RTOL, (where, worst)
    return worst


def _oracle_sample(res, d, x, variant, iters, percent, c, where, limit=None):
    """One unit against the FP64 oracle following the engine, every iteration
    up to limit (all when None)."""
    rows = x.shape[0] if variant == po.SMP else 1
    ora = oracle_following(d, x, variant, iters, percent, res.index[c:c + rows], res.counts[c:c + rows],
                           iter_limit=limit)
    errs = []
    for ch in range(x.shape[0]):
        err = compare_rows(res.params[c + ch], res.index[c + ch], res.counts[c + ch], ora.params[ch],
                           ora.index[ch], ora.counts[ch], ora.followed[0], limited=limit is not None)
        assert_parity(err, f"{where} channel {ch}", min_compared=min(limit or iters, int(res.counts[c + ch])))
        errs.append(err)
    n = int(ora.counts[0])
    assert energy_rel(res.energies[c if variant == po.SMP else 0], ora.energies[0], n) <= ENERGY_RTOL, where
    print(f"{where}: {n} atoms compared, near-ties followed {int(ora.followed[0].sum())}, "
          f"worst amplitude {max(e['amp_rel'] for e in errs):.2e}", flush=True)


def test_default_batch_full_size():
    _announce(f"full size: {CH} ch x {EPOCHS} epochs x {N} samples @{FS:g} Hz, {ITERS} iterations")
    x = synth.batched_channels(items=CH * EPOCHS, N=N, fs=FS)
    specs = [E.ItemSpec(CH, N, FS, E.SMP, ITERS, 99.0, 0.01) for _ in range(EPOCHS)]
    plan = E.Plan(specs)
    plan.run(x.ravel())
    d = po.Dictionary(FS, N, 0.01)
    worst = 0.0
    for e in range(EPOCHS):
        res = plan.result(e)
        assert np.all(res.counts == ITERS), (e, res.counts)
        worst = max(worst, _check_properties(res, d, ITERS, CH, f"epoch {e}"))
    print(f"energy-drop identity over {CH * EPOCHS} units: worst {worst:.2e}", flush=True)

    for e, c, limit in [(0, 0, 30), (EPOCHS // 2, CH // 2, 30), (EPOCHS - 1, CH - 2, 30), (EPOCHS - 1, CH - 1, None)]:
        res = plan.result(e)
        _oracle_sample(res, d, x[e * CH + c][None, :], po.SMP, ITERS, 99.0, c, f"epoch {e} channel {c}", limit)


def test_mmp3_full_size():
    """MMP3 at its full size: 32 channels x 10240 samples at 512 Hz, 12 shared
    atoms, energyError 0.005, 200 iterations."""
    _announce("full size: MMP3 32 ch x 10240 samples @512 Hz, eps 0.005, 200 iterations")
    x = synth.shared_multichannel()
    spec = E.ItemSpec(32, 10240, 512.0, E.MMP3, 200, 99.0, 0.005)
    plan = E.Plan([spec])
    plan.run(x.ravel())
    res = plan.result(0)
    d = po.Dictionary(512.0, 10240, 0.005)
    worst = _check_properties(res, d, 200, 1, "mmp3")
    print(f"energy-drop identity over 200 iterations x 32 channels: worst {worst:.2e}", flush=True)
    n = int(res.counts[0])
    assert np.all(res.counts == n)
    assert np.all(res.index[:, :n] == res.index[:1, :n])


def test_mmp3_config3_dictionary_oracle():
    """MMP3 on config 3's own dictionary (512 Hz, 10240 samples, energyError
    0.005: 62 scales, FFTs up to 262,144 points), the first 4 channels, 3
    iterations, against the oracle at every iteration."""
    _announce("MMP3 4 ch x 10240 @512 Hz, eps 0.005, 3 iterations vs oracle")
    x = synth.shared_multichannel()[:4]
    plan = E.Plan([E.ItemSpec(4, 10240, 512.0, E.MMP3, 3, 99.0, 0.005)])
    plan.run(x.ravel())
    res = plan.result(0)
    d = po.Dictionary(512.0, 10240, 0.005)
    _check_properties(res, d, 3, 1, "mmp3 4 ch")
    _oracle_sample(res, d, x, po.MMP3, 3, 99.0, 0, "mmp3 4 ch x 10240")


def test_uneven_batch_full_size():
    """The uneven batch at its full size: 512 signals of 1024..16384 samples at
    256 Hz, 99% / at most 100 iterations, one plan."""
    _announce("full size: uneven batch, 512 signals of 1024..16384 samples @256 Hz, 100 iterations")
    xs = synth.uneven_batch()
    specs = [E.ItemSpec(1, len(s), 256.0, E.SMP, 100, 99.0, 0.01) for s in xs]
    plan = E.Plan(specs)
    plan.run(np.concatenate(xs))
    dicts = {}
    worst = 0.0
    for i, s in enumerate(xs):
        d = dicts.setdefault(len(s), po.Dictionary(256.0, len(s), 0.01))
        worst = max(worst, _check_properties(plan.result(i), d, 100, 1, f"signal {i}"))
    print(f"energy-drop identity over {len(xs)} signals: worst {worst:.2e}", flush=True)
    order = np.argsort([len(s) for s in xs])
    for i, limit in [(order[0], None), (order[1], None), (order[len(xs) // 2], 30)]:
        i = int(i)
        res = plan.result(i)
        _oracle_sample(res, dicts[len(xs[i])], xs[i][None, :], po.SMP, 100, 99.0, 0,
                       f"signal {i} ({len(xs[i])} samples)", limit)
