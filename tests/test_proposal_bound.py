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
# the device code (k_mark_dirty, k_rows).  A product is P = |u|^2 h(phi)^2,
# u = sqrt(K) v linear in the residual, h = (1 + E_j cos 2 phi)^(-1/2)
# (gabor.cpp:62-70), E_j = exp(-2 pi (j alpha)^2), j counted from the nearer
# end of the spectrum.  For a change |du| <= r of u (r = |a| times the atom
# overlap):
#   (A) sqrt P_new <= (|u_old| + r) / sqrt(1 - E_j)       (1 + E cos 2 phi >= 1 - E)
#   (B) sqrt P_new <= sqrt P_old + L(E_j) r,  L(E) = max |grad (|u| h)|
#       = sqrt(32/27) / (1 - E^2) for E >= 1/3, else 1 / sqrt(1 - E)
# Bins with E_j <= 1e-4 (j >= JB, and the real bins 0 and M) use (B) with
# L(E_JB); bins 1 <= j < JB the smaller of (A) with a per-cell maximum of
# |u|^2 and (B) with L(E_1).  (B) is exact when nothing changed.

def e_of(alpha, j):
    return np.exp(-2 * np.pi * (j * alpha) ** 2)


def lipschitz(E):
    return np.sqrt(32.0 / 27.0) / (1 - E * E) if E >= 1.0 / 3.0 else 1.0 / np.sqrt(1 - E)


def scale_factors(s, fs, Nfft):
    """(JB, gH, gL, lL) of a scale, as the engine's plan sets them."""
    alpha = s * fs / Nfft
    M = Nfft // 2
    JB = int(min(M + 1, np.ceil(np.sqrt(np.log(1e4) / (2 * np.pi)) / alpha)))
    eb = e_of(alpha, JB)
    e1 = e_of(alpha, 1) if M >= 2 else 0.0
    return JB, lipschitz(eb), 1 / np.sqrt(1 - e1), lipschitz(e1)


def _pg(sig2, fs, d):
    """sum_m exp(-pi sig2 (d - m fs)^2): the sampled Gaussian's periodic spectrum."""
    return sum(np.exp(-np.pi * sig2 * (d - m * fs) ** 2) for m in range(-3, 4))


def overlap_rho(s, fs, N, Nfft, JB, t0, t1, params):
    """(all bins, bins 1 <= j < JB) bounds on |u| of the unit atom `params`
    at scale s, time in [t0, t1] (k_mark_dirty).  The untruncated discrete
    overlap of two Gabor atoms is n_A n I / 2 |PG(fA - f) + PG(fA + f)|
    (Poisson summation, I the envelope integral); the low bins lie within f_L
    of 0 and of fs/2.  The samples that the signal's ends, the window (3 s)
    and the atom's support (3 sA) cut off add at most I / sg_n times the
    Gaussian tail sums beyond the cuts (sg_n the width in samples of the two
    envelopes' product, centred on c), taken at the region's end columns; the
    envelope bound n_A n I PG(0) caps both."""
    modulus, amp, sA, fA, tA = params[0], params[1], params[3] / fs, params[4] * fs / 2, params[2] / fs
    nA = np.sqrt(fs) * abs(amp) / abs(modulus)
    dt = t0 - tA if tA < t0 else (tA - t1 if tA > t1 else 0.0)
    v = s * s + sA * sA
    sig2 = s * s * sA * sA / v
    I = s * sA / np.sqrt(v) * np.exp(-np.pi * dt * dt / v)
    n = np.sqrt(2 * np.sqrt(2) / s)
    base = 1.01 * nA * n * I
    sgn2 = sig2 * fs * fs

    def tail(D):
        return np.exp(-np.pi * D * D / sgn2) / (1 - np.exp(-2 * np.pi * D / sgn2)) if D > 0 else np.inf

    loA = max(0.0, (tA - 3 * sA) * fs + 0.5)
    hiA = min(N - 1.0, (tA + 3 * sA) * fs - 0.5)
    cut_lo = cut_hi = 0.0
    for t in (t0, t1):
        c = fs * (t * sA * sA + tA * s * s) / v
        cut_lo = max(cut_lo, tail(c - max(loA, (t - 3 * s) * fs + 0.5) + 1))
        cut_hi = max(cut_hi, tail(min(hiA, (t + 3 * s) * fs - 0.5) + 1 - c))
    cut = (cut_lo + cut_hi) / np.sqrt(sgn2)
    env = base * _pg(sig2, fs, 0.0)
    fN = fs / 2
    fL = max(JB - 1, 0) * fs / Nfft
    d_diff = min(max(0.0, fA - fL), max(0.0, fN - fL - fA))
    d_sum = min(fA, fs - fA - fL, fN - fA, fN - fL + fA)
    if np.isfinite(cut):
        hi = min(env, base * (0.5 * (_pg(sig2, fs, 0.0) + _pg(sig2, fs, min(fA, fN - fA))) + cut))
        lo = min(env, base * (0.5 * (_pg(sig2, fs, d_diff) + _pg(sig2, fs, max(0.0, d_sum))) + cut))
    else:
        hi = lo = env
    return min(1.0, hi + 1e-9), min(1.0, lo + 1e-9)


def region_columns(d, k, b):
    """Time indices of region (scale k, 256-sample block b): columns whose
    centre sample t (N-1)/(T-1) falls in the block."""
    T, N = d.scales[k].T, d.N
    first = lambda bb: (bb * 256 * (T - 1) + (N - 2)) // (N - 1)
    return min(first(b), T), min(first(b + 1) - 1, T - 1)


def low_seeds(d, k, x):
    """Per column: max over bins 1 <= j < JB of |u|^2 = K |v|^2, the product at
    E = 0 (the kernel's num * K')."""
    sc = d.scales[k]
    M = sc.Nfft // 2
    JB = scale_factors(sc.s, d.fs, sc.Nfft)[0]
    kb = np.arange(sc.F)
    j = np.where(4 * kb > sc.Nfft, M - kb, kb)
    low = (j >= 1) & (j < JB)
    out = np.zeros(sc.T)
    if low.any():
        K = d.fs * 2 * np.sqrt(2) / sc.s
        for t in range(sc.T):
            out[t] = K * float(np.max(np.abs(d.column(k, x, t)[low]) ** 2))
    return out


def region_bounds(d, k, tlo, thi, P_old, seeds, atoms):
    """(bound without the E factors, bound) on the region's new maximum,
    without the Cauchy-Schwarz cap: P_old [F][T] and seeds [T] at the last
    evaluation, atoms the parameters subtracted since."""
    sc = d.scales[k]
    t0, t1 = tlo * d.tMax / (sc.T - 1), thi * d.tMax / (sc.T - 1)
    rmax = float(np.max(P_old[:, tlo:thi + 1]))
    seed = float(np.max(seeds[tlo:thi + 1]))
    JB, gH, gL, lL = scale_factors(sc.s, d.fs, sc.Nfft)
    rh = [overlap_rho(sc.s, d.fs, d.N, sc.Nfft, JB, t0, t1, p) for p in atoms]
    racc_hi = sum(abs(p[0]) * r[0] for p, r in zip(atoms, rh))
    racc_lo = sum(abs(p[0]) * r[1] for p, r in zip(atoms, rh))
    old = (np.sqrt(rmax) + sum(abs(p[0]) * _old_rho(sc.s, d.fs, t0, t1, p) for p in atoms)) ** 2 * 1.0001
    hi = (np.sqrt(rmax) + racc_hi * gH) ** 2
    lo = min((np.sqrt(seed) + racc_lo) ** 2 * gL * gL, (np.sqrt(rmax) + racc_lo * lL) ** 2)
    return old, max(hi, lo) * 1.0001


def _old_rho(s, fs, t0, t1, params):
    """The overlap bound the proposal used before the E factors and the
    clipping rule (no counterpart in the device code now)."""
    modulus, amp, sA, fA, tA = params[0], params[1], params[3] / fs, params[4] * fs / 2, params[2] / fs
    nA = np.sqrt(fs) * abs(amp) / abs(modulus)
    dt = t0 - tA if tA < t0 else (tA - t1 if tA > t1 else 0.0)
    v = s * s + sA * sA
    overlap = s * sA / np.sqrt(v) * np.exp(-np.pi * dt * dt / v)
    mirror = np.exp(-np.pi * fA * fA * s * s * sA * sA / v)
    n = np.sqrt(2 * np.sqrt(2) / s)
    alias = 2 * (np.exp(-np.pi * (s * fs) ** 2 / 8) + np.exp(-np.pi * (sA * fs) ** 2 / 8))
    return min(1.0, 1.01 * 0.5 * nA * n * overlap * (1 + mirror) + alias)


def _atom(d, k, f, t, amp, phase):
    """An atom as the loop subtracts it: parameters from the map value of a
    dictionary waveform of amplitude amp and phase `phase` at (k, f, t)
    (gabor.cpp:111-131), and the waveform those parameters give
    (gabor.cpp:143-187; clipped at the edges its amplitude is the map's)."""
    sc = d.scales[k]
    t0 = t * d.tMax / (sc.T - 1)
    w, _ = po.subtract(d.fs, [0.0, amp, t0 * d.fs, sc.s * d.fs, 2.0 * f / sc.Nfft, phase], np.zeros(d.N))
    v = d.column(k, -w, t)[f]
    p = d.atom_params(k, f, t, np.array([v]))[0]
    w, _ = po.subtract(d.fs, p, np.zeros(d.N))
    return -w, p


def _scenario(seed, fs, N, eps, rng=None):
    """A pink-noise residual and one to three random atoms (any scale, bin
    including 0, 1..3 and the last two, position including both edges, phase
    and amplitude) subtracted from it: (dictionary, products before, residual
    after, atom parameters)."""
    rng = rng if rng is not None else np.random.default_rng(100 + seed)
    d = po.Dictionary(fs, N, eps)
    x = synth.pink_noise(rng, N) * 2.0
    L = len(d.scales)
    old_P = [(d.fresh_products(k, x), low_seeds(d, k, x)) for k in range(L)]
    atoms = []
    for _ in range(int(rng.integers(1, 4))):
        k = int(rng.integers(0, L))
        sc = d.scales[k]
        f = int(rng.choice([0, 1, 2, 3, int(rng.integers(0, sc.F)), sc.F - 2, sc.F - 1]))
        f = max(0, min(sc.F - 1, f))
        t = int(rng.choice([0, 1, sc.T - 1, int(rng.integers(0, sc.T))]))
        w, p = _atom(d, k, f, t, float(rng.uniform(0.5, 4.0)) * np.sqrt(sc.s), float(rng.uniform(0, 2 * np.pi)))
        x = x - w
        atoms.append(p)
    return d, old_P, x, atoms


def _regions(d, old_P, x, atoms):
    """(scale, block, new max, old bound, new bound) of every region."""
    out = []
    for k in range(len(d.scales)):
        P_new = d.fresh_products(k, x)
        for b in range((d.N + 255) // 256):
            tlo, thi = region_columns(d, k, b)
            if tlo > thi:
                continue
            old, new = region_bounds(d, k, tlo, thi, old_P[k][0], old_P[k][1], atoms)
            out.append((k, b, float(np.max(P_new[:, tlo:thi + 1])), old, new))
    return out


def test_region_bound_covers_low_bin_phase_turn():
    """Three atoms subtracted next to the signal's start (seed 64): at bin 1 of
    scale 18 a subtraction turns the phase and the new maximum exceeds the
    overlap bound without the E factors by 9%; the bound with them holds for
    every region."""
    d, old_P, x, atoms = _scenario(64, 64.0, 640, 0.02)
    regs = _regions(d, old_P, x, atoms)
    missed = [(k, b, m / old) for k, b, m, old, new in regs if m > old]
    assert any(k == 18 and r > 1.05 for k, b, r in missed), missed
    assert all(m <= new for k, b, m, old, new in regs)


@pytest.mark.parametrize("seed", range(12))
def test_region_bound_holds_on_random_subtractions(seed):
    """Every region's new maximum stays below the bound, at two rates."""
    rng = np.random.default_rng(100 + seed)
    worst = 0.0
    for fs, N, eps in [(64.0, 640, 0.02), (128.0, 512, 0.05)]:
        d, old_P, x, atoms = _scenario(seed, fs, N, eps, rng)
        for k, b, m, old, new in _regions(d, old_P, x, atoms):
            assert m <= new, (seed, fs, k, b, m, new)
            worst = max(worst, m / new)
    print("worst new max / bound", worst)


# ------------------------------------------------------------ Cauchy-Schwarz cover
#
# k_rows bounds a dirty region's products by csnorm times the largest
# residual energy under one of its columns' windows.  Where the columns are
# many it takes windows of Wmax + S samples at starts S = Wmax / 2 apart
# between the first and the last column's start instead, capped by the
# energy under the union of the windows: restated here, that cover must never
# fall below the exact per-column maximum (a gap in the starts would).

def _cs_cover(d, k, tlo, thi, pre):
    sc = d.scales[k]
    N = d.N

    def win(t):
        tc = t * d.tMax / (sc.T - 1)
        wl, wr = np.rint((tc - 3 * sc.s) * d.fs), np.rint((tc + 3 * sc.s) * d.fs)
        return int(max(wl, 0)), int(min(wr, N - 1))

    wmax = max(win(t)[1] - win(t)[0] + 1 for t in range(sc.T))
    il0, il1 = win(tlo)[0], win(thi)[0]
    S = max(1, wmax // 2)
    ncol = thi - tlo + 1
    if ncol > 16 and (il1 - il0) // S + 1 < ncol:
        cover = max(pre[min(i + wmax + S, N)] - pre[i] for i in range(il0, il1 + 1, S))
        return min(cover, pre[win(thi)[1] + 1] - pre[il0])
    return max(pre[win(t)[1] + 1] - pre[win(t)[0]] for t in range(tlo, thi + 1))


@pytest.mark.parametrize("fs,N", [(256.0, 1500), (128.0, 777)])
def test_region_cs_cover(fs, N):
    d = po.Dictionary(fs, N, 0.01)
    rng = np.random.default_rng(int(N))
    for trial in range(3):
        x = synth.pink_noise(rng, N) * rng.uniform(0.5, 2.0, N) ** 3   # uneven energy
        pre = np.concatenate([[0.0], np.cumsum(x * x)])
        for k, sc in enumerate(d.scales):
            for b in range((N + 255) // 256):
                tlo, thi = region_columns(d, k, b)
                if tlo > thi:
                    continue
                wins = []
                for t in range(tlo, thi + 1):
                    tc = t * d.tMax / (sc.T - 1)
                    wins.append((int(max(np.rint((tc - 3 * sc.s) * fs), 0)),
                                 int(min(np.rint((tc + 3 * sc.s) * fs), N - 1))))
                exact = max(pre[r + 1] - pre[l] for l, r in wins)
                union = pre[max(r for _, r in wins) + 1] - pre[min(l for l, _ in wins)]
                cover = _cs_cover(d, k, tlo, thi, pre)
                assert exact <= cover * (1 + 1e-12), (k, b, exact, cover)
                assert cover <= union * (1 + 1e-12) + 1e-12, (k, b, cover, union)
