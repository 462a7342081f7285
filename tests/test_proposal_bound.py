"""The proposal's certified bound P <= c * (residual energy under the
column's window), c = empi_proposal_bound_factor(s, fs, Nfft), checked on the
FP64 oracle's own products (gabor.cpp:44-70 via the oracle's column and
product): on noise, on pink noise, and on the Cauchy-Schwarz equality case
where the residual is the cell's own atom (so the bound must also be nearly
tight there).  Host-only: no device call."""
import numpy as np
import pytest

import empi_amd as E
import pyoracle as po
from empi_amd import synth


def _window(d, k, t):
    sc = d.scales[k]
    t0 = t * d.tMax / (sc.T - 1)
    hw = 3.0 * sc.s
    iL = max(0, int(np.rint((t0 - hw) * d.fs)))
    iR = min(d.N - 1, int(np.rint((t0 + hw) * d.fs)))
    return t0, iL, iR


def _check_cells(d, x, k, ts, fs_bins):
    sc = d.scales[k]
    c = E.proposal_bound_factor(sc.s, d.fs, sc.Nfft)
    worst = 0.0
    for t in ts:
        _, iL, iR = _window(d, k, t)
        ew = float(np.sum(x[iL:iR + 1] ** 2))
        col = d.column(k, x, t)
        for f in fs_bins:
            p = d.product(k, f, t, np.array([col[f]]))
            assert p <= c * ew * (1 + 1e-12), (k, t, f, p, c * ew)
            worst = max(worst, p / (c * ew))
    return worst


@pytest.mark.parametrize("fs,N", [(128.0, 1280), (256.0, 5120)])
def test_bound_holds_on_noise(fs, N):
    d = po.Dictionary(fs, N, 0.01)
    rng = np.random.default_rng(7)
    white = rng.standard_normal(N)
    pink = synth.pink_noise(np.random.default_rng(8), N)
    for k, sc in enumerate(d.scales):
        if 6 * sc.s * fs >= N / 4:        # scales the proposal always evaluates
            continue
        ts = np.unique(np.linspace(0, sc.T - 1, 5).astype(int))
        bins = np.unique(np.concatenate([np.arange(min(sc.F, 6)), np.arange(max(0, sc.F - 4), sc.F),
                                         np.linspace(0, sc.F - 1, 12).astype(int)]))
        for x in (white, pink):
            _check_cells(d, x, k, ts, bins)


@pytest.mark.parametrize("fs,N", [(128.0, 1280), (256.0, 5120)])
def test_bound_tight_on_matched_atoms(fs, N):
    """Residual = the cell's own atom (any phase, low, mid and Nyquist-side
    bins): Cauchy-Schwarz is an equality up to the discrete norm, so
    P / (c E) must stay <= 1 and reach close to 1."""
    d = po.Dictionary(fs, N, 0.01)
    for k, sc in enumerate(d.scales):
        if 6 * sc.s * fs >= N / 4:
            continue
        best = 0.0
        t = sc.T // 2
        t0, iL, iR = _window(d, k, t)
        M = sc.Nfft // 2
        for f in sorted({0, 1, 2, 3, sc.F // 3, M - 2, M - 1, M} & set(range(sc.F))):
            for phase in np.linspace(0, np.pi, 9):
                params = [0.0, 1.0, t0 * fs, sc.s * fs, 2.0 * f / sc.Nfft, phase]
                atom, _ = po.subtract(fs, params, np.zeros(N))
                atom = -atom
                if not np.any(atom[iL:iR + 1]):
                    continue
                x = np.zeros(N)
                x[iL:iR + 1] = atom[iL:iR + 1]
                best = max(best, _check_cells(d, x, k, [t], [f]))
        # beyond 6 samples per scale the aliasing terms vanish: c is exact to 1e-4
        assert best > (0.999 if sc.s * fs >= 6 else 0.1), (k, sc.s * fs, best)
