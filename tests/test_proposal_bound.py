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


# ------------------------------------------------------------ region bound
#
# The proposal's bound on a dirty region's new maximum, restated in FP64 from
# the device code (k_mark_dirty, k_rows): for every bin j >= 1 of every cell,
# with P = K |v|^2 / (1 + E_j cos 2 phi) (gabor.cpp:62-70), E_j = exp(-2 pi (j
# alpha)^2) and j counted from the nearer end of the spectrum,
#   sqrt(P_new) <= sqrt(P_old (1 + E_j) / (1 - E_j)) + sum |a| rho / sqrt(1 - E_j),
# since 1 - E <= 1 + E cos 2 phi <= 1 + E and sqrt(K)|v| is a seminorm of the
# residual; bins 0 and M are real-valued with a fixed denominator.  Bins with
# E_j > 1e-4 (j < JB) use a per-cell maximum of P (1 + E)/(1 - E) and the
# factor of bin 1; the others the region maximum.

def e_of(alpha, j):
    return np.exp(-2 * np.pi * (j * alpha) ** 2)


def scale_factors(s, fs, Nfft):
    """(JB, fH, gH, gL) of a scale, as the engine's plan sets them."""
    alpha = s * fs / Nfft
    M = Nfft // 2
    JB = int(min(M + 1, np.ceil(np.sqrt(np.log(1e4) / (2 * np.pi)) / alpha)))
    eb = e_of(alpha, JB)
    e1 = e_of(alpha, 1) if M >= 2 else 0.0
    return JB, (1 + eb) / (1 - eb), 1 / np.sqrt(1 - eb), 1 / np.sqrt(1 - e1)


def overlap_rho(s, fs, t0, t1, params, clipped):
    """Bound on the overlap of the unit atom `params` with any unit real atom
    of scale s and time in [t0, t1] (k_mark_dirty), at E = 0."""
    modulus, amp, sA, fA, tA = params[0], params[1], params[3] / fs, params[4] * fs / 2, params[2] / fs
    nA = np.sqrt(fs) * abs(amp) / abs(modulus)
    dt = t0 - tA if tA < t0 else (tA - t1 if tA > t1 else 0.0)
    v = s * s + sA * sA
    overlap = s * sA / np.sqrt(v) * np.exp(-np.pi * dt * dt / v)
    mirror = 1.0 if clipped else np.exp(-np.pi * fA * fA * s * s * sA * sA / v)
    n = np.sqrt(2 * np.sqrt(2) / s)
    alias = 2 * (np.exp(-np.pi * (s * fs) ** 2 / 8) + np.exp(-np.pi * (sA * fs) ** 2 / 8))
    return min(1.0, 1.01 * 0.5 * nA * n * overlap * (1 + mirror) + alias)


def region_columns(d, k, b):
    """Time indices of region (scale k, 256-sample block b): columns whose
    centre sample t (N-1)/(T-1) falls in the block."""
    T, N = d.scales[k].T, d.N
    first = lambda bb: (bb * 256 * (T - 1) + (N - 2)) // (N - 1)
    return min(first(b), T), min(first(b + 1) - 1, T - 1)


def _clipped(d, k, tlo, thi, params):
    s = d.scales[k].s
    t0, t1 = tlo * d.tMax / (d.scales[k].T - 1), thi * d.tMax / (d.scales[k].T - 1)
    sA, tA = params[3] / d.fs, params[2] / d.fs
    return (np.rint((t0 - 3 * s) * d.fs) < 0 or np.rint((t1 + 3 * s) * d.fs) > d.N - 1 or
            np.rint((tA - 3 * sA) * d.fs) < 0 or np.rint((tA + 3 * sA) * d.fs) > d.N - 1)


def low_inflated(d, k, P):
    """Per column: max over bins 1 <= j < JB of P (1 + E_j) / (1 - E_j)."""
    sc = d.scales[k]
    alpha = sc.s * d.fs / sc.Nfft
    M = sc.Nfft // 2
    JB = scale_factors(sc.s, d.fs, sc.Nfft)[0]
    kb = np.arange(sc.F)
    j = np.where(4 * kb > sc.Nfft, M - kb, kb)
    low = (j >= 1) & (j < JB)
    if not low.any():
        return np.zeros(P.shape[1])
    E = e_of(alpha, j[low])
    return np.max(P[low] * ((1 + E) / (1 - E))[:, None], axis=0)


def region_bounds(d, k, tlo, thi, P_old, atoms):
    """(old bound, new bound) on the region's new maximum without the
    Cauchy-Schwarz cap: P_old [F][T] the products at the last evaluation,
    atoms the parameters subtracted since."""
    sc = d.scales[k]
    t0, t1 = tlo * d.tMax / (sc.T - 1), thi * d.tMax / (sc.T - 1)
    rmax = float(np.max(P_old[:, tlo:thi + 1]))
    rlow = float(np.max(low_inflated(d, k, P_old)[tlo:thi + 1]))
    racc = sum(abs(p[0]) * overlap_rho(sc.s, d.fs, t0, t1, p, _clipped(d, k, tlo, thi, p)) for p in atoms)
    racc_old = sum(abs(p[0]) * overlap_rho(sc.s, d.fs, t0, t1, p, False) for p in atoms)
    _, fH, gH, gL = scale_factors(sc.s, d.fs, sc.Nfft)
    old = (np.sqrt(rmax) + racc_old) ** 2 * 1.0001
    new = max((np.sqrt(rmax * fH) + racc * gH) ** 2, (np.sqrt(rlow) + racc * gL) ** 2) * 1.0001
    return old, new


def _atom(d, k, f, t, amp, phase):
    """A dictionary atom's waveform and its book parameters (its own map value)."""
    sc = d.scales[k]
    t0 = t * d.tMax / (sc.T - 1)
    w, _ = po.subtract(d.fs, [0.0, amp, t0 * d.fs, sc.s * d.fs, 2.0 * f / sc.Nfft, phase], np.zeros(d.N))
    w = -w
    v = d.column(k, w, t)[f]
    return w, d.atom_params(k, f, t, np.array([v]))[0]


def test_region_bound_covers_low_bin_phase_turn():
    """A subtraction that turns a low bin's phase from cosine to sine raises its
    product by up to (1 + E)/(1 - E) at unchanged |v|: the cell holds a
    cosine-phase atom A plus a sine-phase atom of 3x the amplitude at bin 1.
    The overlap bound without the E factors misses the new maximum; the
    bound with them holds."""
    fs, N = 128.0, 1280
    d = po.Dictionary(fs, N, 0.01)
    k = next(i for i, sc in enumerate(d.scales) if sc.s * fs >= 20 and sc.T > 8)
    sc = d.scales[k]
    t = sc.T // 2
    A, pA = _atom(d, k, 1, t, 1.0, 0.0)
    S, _ = _atom(d, k, 1, t, 3.0, np.pi / 2)
    x = A + S
    b = int(np.rint(t * (N - 1) / (sc.T - 1))) // 256
    tlo, thi = region_columns(d, k, b)
    assert tlo <= t <= thi
    P_old = d.fresh_products(k, x)
    P_new = d.fresh_products(k, x - A)
    new_max = float(np.max(P_new[:, tlo:thi + 1]))
    old, new = region_bounds(d, k, tlo, thi, P_old, [pA])
    assert old < new_max, (old, new_max)        # the bound without the E factors fails here
    assert new_max <= new, (new_max, new)


@pytest.mark.parametrize("seed", range(6))
def test_region_bound_holds_on_random_subtractions(seed):
    """Random pink-noise residuals, random atoms (any scale, position, bin,
    phase and amplitude; the signal's edges included) subtracted one to three
    at a time: every region of every scale that the reference rule marks
    dirty keeps its new maximum below the bound."""
    rng = np.random.default_rng(100 + seed)
    fs, N = 64.0, 640
    d = po.Dictionary(fs, N, 0.02)
    x = synth.pink_noise(rng, N) * 2.0
    L = len(d.scales)
    old_P = [d.fresh_products(k, x) for k in range(L)]
    atoms = []
    for _ in range(int(rng.integers(1, 4))):
        k = int(rng.integers(0, L))
        sc = d.scales[k]
        f = int(rng.choice([1, 2, 3, int(rng.integers(0, sc.F)), sc.F - 2]))
        f = max(0, min(sc.F - 1, f))
        t = int(rng.choice([0, 1, sc.T - 1, int(rng.integers(0, sc.T))]))
        w, p = _atom(d, k, f, t, float(rng.uniform(0.5, 4.0)) * np.sqrt(sc.s), float(rng.uniform(0, 2 * np.pi)))
        x = x - w
        atoms.append(p)
    B = (N + 255) // 256
    worst = 0.0
    for k in range(L):
        sc = d.scales[k]
        P_new = d.fresh_products(k, x)
        for b in range(B):
            tlo, thi = region_columns(d, k, b)
            if tlo > thi:
                continue
            t0, t1 = tlo * d.tMax / (sc.T - 1), thi * d.tMax / (sc.T - 1)
            if not any(abs(tt - p[2] / fs) <= 3 * (sc.s + p[3] / fs) + d.tMax / (sc.T - 1)
                       for p in atoms for tt in (t0, t1, np.clip(p[2] / fs, t0, t1))):
                continue
            _, new = region_bounds(d, k, tlo, thi, old_P[k], atoms)
            new_max = float(np.max(P_new[:, tlo:thi + 1]))
            worst = max(worst, new_max / new)
            assert new_max <= new, (seed, k, b, new_max, new)
    print("worst new max / bound", worst)
