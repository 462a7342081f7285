"""FFT, window and map columns of the oracle (inc/fftw.hpp:120-127,
src/gabor.cpp:31-40, 279-303) against brute force and closed forms."""
import numpy as np
import pytest

import pyoracle as po
from empi_amd import synth


def brute_dft(x):
    n = len(x)
    k = np.arange(n // 2 + 1)[:, None]
    t = np.arange(n)[None, :]
    ang = -2.0 * np.pi * ((k * t) % n) / n
    return (x[None, :] * (np.cos(ang) + 1j * np.sin(ang))).sum(axis=1)


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 256, 1024])
def test_rfft_brute_force(n):
    x = np.random.default_rng(n).standard_normal(n)
    X = po.rfft(x)
    ref = brute_dft(x)
    assert np.max(np.abs(X - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [4096, 65536, 262144])
def test_rfft_numpy(n):
    x = np.random.default_rng(n).standard_normal(n)
    X = po.rfft(x)
    assert np.max(np.abs(X - np.fft.rfft(x))) <= 1e-13 * np.max(np.abs(X))


def test_rfft_pure_tone_single_bin():
    n, k0 = 512, 37
    x = np.cos(2 * np.pi * k0 * np.arange(n) / n)
    X = po.rfft(x)
    assert abs(X[k0] - n / 2) < 1e-10
    X[k0] = 0
    assert np.max(np.abs(X)) < 1e-10


def gabor_inner(d, k, x, t):
    """sum_n x[n] exp(-pi (n/fs - t0)^2 / s^2) e^{-2 pi i f (n/fs - t0)} / fs over the window."""
    sc = d.scales[k]
    fs = d.fs
    t0 = t * d.tMax / (sc.T - 1)
    iL = max(0, int(np.rint((t0 - 3 * sc.s) * fs)))
    iR = min(d.N - 1, int(np.rint((t0 + 3 * sc.s) * fs)))
    n = np.arange(iL, iR + 1)
    tt = n / fs - t0
    g = np.exp(-np.pi * tt ** 2 / sc.s ** 2)
    f = np.arange(sc.F) * fs / sc.Nfft
    return ((x[n] * g)[None, :] * np.exp(-2j * np.pi * f[:, None] * tt[None, :])).sum(axis=1) / fs


@pytest.mark.parametrize("k,t", [(0, 0), (0, 7995), (3, 100), (10, 500), (15, 0), (20, 160), (25, 30),
                                 (33, 0), (33, 7), (33, 13)])
def test_column_is_gabor_inner_product(k, t):
    x = synth.single_channel().astype(np.float64)
    d = po.Dictionary(128.0, 1280, 0.01)
    v = d.column(k, x, t)
    ref = gabor_inner(d, k, x, t)
    assert np.max(np.abs(v - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_column_linear():
    d = po.Dictionary(128.0, 1280, 0.01)
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal(1280), rng.standard_normal(1280)
    for k, t in [(4, 1000), (22, 50)]:
        lhs = d.column(k, 2.0 * a - 3.0 * b, t)
        rhs = 2.0 * d.column(k, a, t) - 3.0 * d.column(k, b, t)
        assert np.max(np.abs(lhs - rhs)) <= 1e-12 * np.max(np.abs(lhs))
