"""Seeded synthetic signals for the five benchmark workloads.

Atoms follow the reference's waveform convention
A * exp(-pi (t-u)^2 / s^2) * cos(2 pi f (t-u) + phi) at t = n/fs
(src/gabor.cpp:158-163), untruncated, with u ~ U[0, (N-1)/fs].
Pink noise: white N(0,1) -> rfft -> bin k >= 1 scaled by k^-1/2, DC = 0 ->
irfft -> unit variance.  Seeds are 1000 * workload + item.
"""
from __future__ import annotations

import numpy as np


def gabor_atom(N, fs, amplitude, scale, freq, phase, position):
    t = np.arange(N) / fs - position
    return amplitude * np.exp(-np.pi * t * t / (scale * scale)) * np.cos(2 * np.pi * freq * t + phase)


def pink_noise(rng, N):
    white = rng.standard_normal(N)
    spec = np.fft.rfft(white)
    k = np.arange(len(spec))
    spec[0] = 0.0
    spec[1:] /= np.sqrt(k[1:])
    x = np.fft.irfft(spec, n=N)
    sd = x.std()
    return x / sd if sd > 0 else x


def _random_atoms(rng, count, N, fs):
    atoms = []
    for _ in range(count):
        s = rng.uniform(0.1, 2.0)
        f = rng.uniform(0.0, 40.0)
        phi = rng.uniform(0.0, 2 * np.pi)
        A = rng.uniform(1.0, 5.0)
        u = rng.uniform(0.0, (N - 1) / fs)
        atoms.append((A, s, f, phi, u))
    return atoms


def single_channel(seed=1001, N=1280, fs=128.0, atoms=5, noise_sigma=0.1):
    """SMP single channel, 1280 samples at 128 Hz: 5 Gabor atoms + white noise 0.1."""
    rng = np.random.default_rng(seed)
    x = np.zeros(N)
    for (A, s, f, phi, u) in _random_atoms(rng, atoms, N, fs):
        x += gabor_atom(N, fs, A, s, f, phi, u)
    x += rng.normal(0.0, noise_sigma, N)
    return x.astype(np.float32)


def pink_with_atoms(seed, N, fs, atoms):
    rng = np.random.default_rng(seed)
    x = pink_noise(rng, N)
    for (A, s, f, phi, u) in _random_atoms(rng, atoms, N, fs):
        x += gabor_atom(N, fs, A, s, f, phi, u)
    return x.astype(np.float32)


def batched_channels(items=1280, N=5120, fs=256.0, atoms=8):
    """SMP batch: channel-epochs of unit-variance pink noise + 8 atoms (seeds 2000+i)."""
    return np.stack([pink_with_atoms(2000 + i, N, fs, atoms) for i in range(items)])


def shared_multichannel(C=32, N=10240, fs=512.0, atoms=12, noise_sigma=0.2, seed=3000):
    """MMP3 workload: 12 atoms share (s, f, u); per-channel A ~ U[0.5,3], phi ~ U[0,2pi)."""
    rng = np.random.default_rng(seed)
    x = np.zeros((C, N))
    for _ in range(atoms):
        s = rng.uniform(0.1, 2.0)
        f = rng.uniform(0.0, 40.0)
        u = rng.uniform(0.0, (N - 1) / fs)
        for c in range(C):
            A = rng.uniform(0.5, 3.0)
            phi = rng.uniform(0.0, 2 * np.pi)
            x[c] += gabor_atom(N, fs, A, s, f, phi, u)
    x += rng.normal(0.0, noise_sigma, (C, N))
    return x.astype(np.float32)


def uneven_batch(items=512, fs=256.0, nmin=1024, nmax=16384, seed0=4000):
    """Uneven lengths U{1024..16384}, pink noise + Poisson(6) atoms (seeds seed0+i)."""
    out = []
    for i in range(items):
        rng = np.random.default_rng(seed0 + i)
        N = int(rng.integers(nmin, nmax + 1))
        count = int(rng.poisson(6))
        x = pink_noise(rng, N)
        for (A, s, f, phi, u) in _random_atoms(rng, count, N, fs):
            x += gabor_atom(N, fs, A, s, f, phi, u)
        out.append(x.astype(np.float32))
    return out
