"""Edge cases of the batched engine against the FP64 oracle (needs a GPU):
dictionary options (maxAtomFrequency clipping, scale bounds), the one-scale
dictionary of a two-sample signal, a single iteration (the last atom is not
subtracted, classes.cpp:57-59), an early energyPercent stop, and one plan
holding every variant and several lengths at once."""
import numpy as np
import pytest

import empi_amd as E
import pyoracle as po
from empi_amd import synth
from test_gpu_parity import _announce, _check_item, _oracle, _run_plan

pytestmark = pytest.mark.gpu


def test_frequency_clip_and_scale_bounds():
    _announce("dictionary options: freqMax 20 Hz, scales 0.05..1 s, 1280 samples @128 Hz")
    x = synth.single_channel(seed=1201)
    spec = E.ItemSpec(1, 1280, 128.0, E.SMP, 25, 99.0, 0.01, scale_min=0.05, scale_max=1.0, freq_max=20.0)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(x[None, :], 128.0, 0.01, po.SMP, 25, 99.0, scale_min=0.05, scale_max=1.0, freq_max=20.0)
    _check_item(res, ora, E.SMP, "clipped dictionary")
    d = po.Dictionary(128.0, 1280, 0.01, 0.05, 1.0, 20.0)
    assert plan.atom_count(0) == d.entries
    # every atom within the clipped band (frequency in units of Nyquist)
    n = int(res.counts[0])
    assert np.all(res.params[0, :n, 4] * 64.0 <= 20.0 + 64.0 / 8)


def test_two_sample_signal_single_scale():
    _announce("two-sample signal: a one-scale dictionary")
    x = np.array([0.75, -1.5], dtype=np.float32)
    spec = E.ItemSpec(1, 2, 16.0, E.SMP, 3, 99.0, 0.01)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    ora = _oracle(x[None, :], 16.0, 0.01, po.SMP, 3, 99.0)
    assert res.status == 0
    _check_item(res, ora, E.SMP, "two samples")


def test_single_iteration_and_early_stop():
    _announce("iterationMax 1 (no subtraction) and energyPercent 50 (early stop)")
    x = synth.pink_with_atoms(1301, 2048, 256.0, 6)
    specs = [E.ItemSpec(1, 2048, 256.0, E.SMP, 1, 99.0, 0.01), E.ItemSpec(1, 2048, 256.0, E.SMP, 60, 50.0, 0.01)]
    plan = _run_plan(specs, [x, x])
    one = plan.result(0)
    assert one.status == 0 and int(one.counts[0]) == 1
    e = one.energies[0]
    assert e[0] > 0 and np.isnan(e[1])          # the only atom is not subtracted
    ora1 = _oracle(x[None, :], 256.0, 0.01, po.SMP, 1, 99.0)
    _check_item(one, ora1, E.SMP, "one iteration")
    early = plan.result(1)
    ora2 = _oracle(x[None, :], 256.0, 0.01, po.SMP, 60, 50.0)
    _check_item(early, ora2, E.SMP, "early stop")
    n = int(early.counts[0])
    assert n < 60
    assert early.energies[0][n] / early.energies[0][0] <= 0.5
    assert early.energies[0][n - 1] / early.energies[0][0] > 0.5


def test_mixed_variants_one_plan():
    _announce("one plan: SMP 2 ch, MMP1/2/3 4 ch, lengths 900..2000, two sampling rates")
    xs, specs, oras = [], [], []
    cases = [(E.SMP, 2, 1500, 256.0), (E.MMP1, 4, 900, 128.0), (E.MMP2, 4, 2000, 256.0), (E.MMP3, 4, 1280, 128.0)]
    for i, (variant, C, N, fs) in enumerate(cases):
        x = synth.shared_multichannel(C=C, N=N, fs=fs, atoms=3, noise_sigma=0.1, seed=1400 + i)
        xs.append(x)
        specs.append(E.ItemSpec(C, N, fs, variant, 12, 99.5, 0.01))
        oras.append(_oracle(x, fs, 0.01, variant, 12, 99.5))
    plan = _run_plan(specs, xs)
    for i, (variant, *_rest) in enumerate(cases):
        res = plan.result(i)
        assert res.status == 0
        _check_item(res, oras[i], variant, f"mixed item {i} (variant {variant})")


def test_full_build_counts_every_entry():
    """A one-iteration run transforms the whole map once and nothing after it
    (the last atom is not subtracted), so the entry counters, which the
    column kernels keep in one copy per XCD and the host sums, must total
    every unit's dictionary size: one per channel of an SMP item, one per
    MMP3 item."""
    _announce("entry counters: one iteration of a mixed SMP / MMP3 batch counts every map entry once")
    rng = np.random.default_rng(77)
    specs = [E.ItemSpec(1, 1280, 128.0, E.SMP, 1, 99.0, 0.01),
             E.ItemSpec(3, 3000, 256.0, E.SMP, 1, 99.0, 0.01),
             E.ItemSpec(1, 20000, 256.0, E.SMP, 1, 99.0, 0.01),
             E.ItemSpec(4, 2000, 256.0, E.MMP3, 1, 99.0, 0.01)]
    plan = _run_plan(specs, [rng.standard_normal((s.channels, s.samples)) for s in specs])
    units = [s.channels if s.variant == E.SMP else 1 for s in specs]
    expected = sum(u * plan.atom_count(i) for i, u in enumerate(units))
    assert plan.entries_evaluated() == expected
