"""MP loop of the oracle (src/classes.cpp:26-87, src/mmp.cpp:23-45):
brute-force argmax, stop rule, energy bookkeeping and variant semantics."""
import numpy as np
import pytest

import pyoracle as po
from empi_amd import synth


def all_products(d, x):
    """Brute force: every (k, f, t) product of a single channel."""
    best = (-1.0, None)
    second = -1.0
    for k, sc in enumerate(d.scales):
        for t in range(sc.T):
            col = d.column(k, x, t)
            for f in range(sc.F):
                p = d.product(k, f, t, [col[f]])
                if p > best[0]:
                    second = max(second, best[0])
                    best = (p, (k, f, t))
                elif p > second:
                    second = p
    return best, second


def test_first_atom_is_brute_force_argmax():
    x = synth.pink_with_atoms(77, 96, 32.0, 2).astype(np.float64)
    d = po.Dictionary(32.0, 96, 0.05)
    (p, idx), second = all_products(d, x)
    r = po.decompose(d, x[None, :], po.SMP, 3, 99.9)
    assert tuple(r.index[0, 0]) == idx
    assert abs(r.gaps[0, 0] - (p - second) / p) < 1e-12


def test_stop_rule_and_energies():
    x = synth.single_channel()
    d = po.Dictionary(128.0, 1280, 0.01)
    r = po.decompose(d, x[None, :].astype(np.float64), po.SMP, 50, 99.0)
    n = int(r.counts[0])
    e = r.energies[0]
    assert e[0] == pytest.approx(np.sum(x.astype(np.float64) ** 2) / 128.0, rel=1e-14)
    assert n < 50
    # classes.cpp:67-70: stop after the first subtraction reaching 1% residual
    assert e[n] / e[0] <= 0.01 and all(e[i] / e[0] > 0.01 for i in range(1, n))
    assert np.all(np.diff(e[: n + 1]) < 0)
    assert np.all(np.isnan(e[n + 1:]))


def test_iteration_max_last_atom_not_subtracted():
    x = synth.single_channel()
    d = po.Dictionary(128.0, 1280, 0.01)
    r = po.decompose(d, x[None, :].astype(np.float64), po.SMP, 4, 99.99)
    assert int(r.counts[0]) == 4
    e = r.energies[0]
    assert not np.isnan(e[3]) and np.isnan(e[4])   # 3 subtractions for 4 atoms


def test_signal_is_empty():
    d = po.Dictionary(128.0, 256, 0.01)
    with pytest.raises(po.OracleError) as ei:
        po.decompose(d, np.zeros((1, 256)), po.SMP, 5, 99.0)
    assert str(ei.value) == "signalIsEmpty"


def test_smp_channels_independent():
    xs = np.stack([synth.pink_with_atoms(10 + c, 512, 128.0, 3) for c in range(3)]).astype(np.float64)
    d = po.Dictionary(128.0, 512, 0.02)
    r = po.decompose(d, xs, po.SMP, 6, 99.0)
    for c in range(3):
        rc = po.decompose(d, xs[c:c + 1], po.SMP, 6, 99.0)
        assert np.array_equal(rc.index[0], r.index[c]) and np.array_equal(rc.params[0], r.params[c])


def test_mmp2_is_sum_decomposition_plus_refit():
    xs = synth.shared_multichannel(C=3, N=512, fs=128.0, atoms=3, seed=5).astype(np.float64)
    d = po.Dictionary(128.0, 512, 0.02)
    r = po.decompose(d, xs, po.MMP2, 5, 99.9)
    rs = po.decompose(d, xs.sum(axis=0, keepdims=True), po.SMP, 5, 99.9)
    n = int(rs.counts[0])
    for c in range(3):
        assert np.array_equal(r.index[c, :n], rs.index[0, :n])
        assert np.array_equal(r.params[c, :n, 2:], rs.params[0, :n, 2:])
        # sequential least-squares refits on channel c
        y = xs[c].copy()
        for i in range(n):
            y, q = po.subtract(128.0, rs.params[0, i], y, fit=True)
            assert np.allclose(q[:2], r.params[c, i, :2], rtol=1e-12, atol=1e-15)


def test_mmp3_product_is_channel_sum():
    xs = synth.shared_multichannel(C=4, N=512, fs=128.0, atoms=2, seed=6).astype(np.float64)
    d = po.Dictionary(128.0, 512, 0.02)
    for k, t, f in [(5, 40, 3), (12, 10, 20)]:
        vals = [d.column(k, xs[c], t)[f] for c in range(4)]
        p = d.product(k, f, t, vals)
        assert abs(p - sum(d.product(k, f, t, [v]) for v in vals)) <= 1e-12 * p


def test_mmp1_shares_phase_across_channels():
    xs = synth.shared_multichannel(C=4, N=512, fs=128.0, atoms=3, seed=7).astype(np.float64)
    d = po.Dictionary(128.0, 512, 0.02)
    r = po.decompose(d, xs, po.MMP1, 5, 99.9)
    for i in range(int(r.counts[0])):
        k, f, t = r.index[0, i]
        fv = f * 128.0 / d.scales[k].Nfft
        if fv > 32.0:          # high-frequency bins skip the constraint
            continue
        ph = r.params[:, i, 5]
        diff = np.angle(np.exp(2j * (ph - ph[0])))   # equal modulo pi
        assert np.max(np.abs(diff)) < 1e-9
