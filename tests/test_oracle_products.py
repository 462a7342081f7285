"""Products, atom parameters, subtraction and the MMP1 constraint of the
oracle (src/gabor.cpp:44-71,111-131,143-187; src/mmp.cpp:9-21) against
closed forms and invariants."""
import numpy as np
import pytest

import pyoracle as po


def atom_signal(d, k, j, i, A, phi):
    sc = d.scales[k]
    u = i * d.tMax / (sc.T - 1)
    f = j * d.fs / sc.Nfft
    t = np.arange(d.N) / d.fs - u
    return A * np.exp(-np.pi * t * t / sc.s ** 2) * np.cos(2 * np.pi * f * t + phi)


def test_dictionary_atom_product_equals_energy():
    """x = A g(s,f,u,phi) with exp(-2 pi s^2 f^2) ~ 0: the product at (s,f,u)
    is fs * energy = sum x^2, amplitude A, phase phi (Cauchy-Schwarz equality)."""
    d = po.Dictionary(128.0, 2048, 0.01)
    for k, j_hz, i_frac, A, phi in [(20, 10.0, 0.5, 3.0, 0.7), (24, 25.0, 0.3, 1.5, -2.0), (16, 40.0, 0.6, 2.0, 2.9)]:
        sc = d.scales[k]
        j = int(round(j_hz * sc.Nfft / d.fs))
        i = int(i_frac * (sc.T - 1))
        x = atom_signal(d, k, j, i, A, phi)
        assert np.exp(-2 * np.pi * (sc.s * j * d.fs / sc.Nfft) ** 2) < 1e-12
        v = d.column(k, x, i)[j]
        P = d.product(k, j, i, [v])
        assert abs(P / np.sum(x * x) - 1) < 1e-6
        p = d.atom_params(k, j, i, [v])[0]
        assert abs(p[1] / A - 1) < 1e-6
        assert abs(np.angle(np.exp(1j * (p[5] - phi)))) < 1e-6
        assert abs(p[2] - i * d.tMax / (sc.T - 1) * d.fs) < 1e-9
        assert abs(p[3] - sc.s * d.fs) < 1e-9 and abs(p[4] - 2 * j / sc.Nfft) < 1e-12
        # MP's first pick on this signal is the atom itself
        r = po.decompose(d, x[None, :], po.SMP, 1, 99.0)
        assert tuple(r.index[0, 0]) == (k, j, i)


@pytest.mark.parametrize("k,i_frac,j_frac", [(8, 0.5, 0.1), (14, 0.4, 0.02), (14, 0.45, 0.7), (22, 0.5, 0.01),
                                             (22, 0.55, 0.3), (25, 0.5, 0.9)])
def test_subtraction_removes_product_energy(k, i_frac, j_frac):
    """Removing amplitude * g_theta (theta = arg v) lowers the energy by P/fs,
    the squared projection on the unit-normalised real atom."""
    d = po.Dictionary(128.0, 2048, 0.01)
    x = np.random.default_rng(k).standard_normal(d.N)
    sc = d.scales[k]
    i = int(i_frac * (sc.T - 1))
    j = max(1, int(j_frac * (sc.F - 1)))
    v = d.column(k, x, i)[j]
    P = d.product(k, j, i, [v])
    p = d.atom_params(k, j, i, [v])[0]
    e0 = np.sum(x * x) / d.fs
    y, _ = po.subtract(d.fs, p, x, fit=False)
    e1 = np.sum(y * y) / d.fs
    assert abs((e0 - e1) / (P / d.fs) - 1) < 1e-6
    assert P <= np.sum(x * x) * (1 + 1e-9)


def test_least_squares_refit_orthogonal():
    """fit=True projects on the atom: residual orthogonal to it (gabor.cpp:165-183)."""
    d = po.Dictionary(128.0, 1280, 0.01)
    x = np.random.default_rng(9).standard_normal(d.N)
    k, i, j = 15, 200, 40
    v = d.column(k, x, i)[j]
    p = d.atom_params(k, j, i, [v])[0]
    y, q = po.subtract(d.fs, p, x, fit=True)
    w = x - y
    assert abs(np.dot(y, w)) < 1e-10 * np.dot(x, x)
    assert abs(q[0] - np.sqrt(np.dot(w, w))) < 1e-9 * abs(q[0])


def test_mmp1_constraint_invariants():
    rng = np.random.default_rng(3)
    for C in (1, 2, 5, 32):
        v = rng.standard_normal(C) + 1j * rng.standard_normal(C)
        out = po.mmp1_constraint(v)
        theta = 0.5 * np.angle(np.sum(v * v))
        # every output lies on the line of angle theta
        assert np.max(np.abs(np.imag(out * np.exp(-1j * theta)))) < 1e-12 * np.max(np.abs(v))
        # and is the projection of its input on that line
        proj = np.real(v * np.exp(-1j * theta)) * np.exp(1j * theta)
        assert np.max(np.abs(out - proj)) < 1e-12 * np.max(np.abs(v))
        # theta maximises the summed squared projection: the top eigenvalue of the 2x2 covariance
        srr, sii, sri = np.sum(v.real ** 2), np.sum(v.imag ** 2), np.sum(v.real * v.imag)
        lam = 0.5 * (srr + sii + np.hypot(srr - sii, 2 * sri))
        assert abs(np.sum(np.abs(out) ** 2) / lam - 1) < 1e-12


def test_mmp1_two_channel_closed_form():
    v = np.array([1.0 + 0j, 1j])          # v1^2 + v2^2 = 0: theta = 0
    out = po.mmp1_constraint(v)
    assert np.allclose(out, [1.0, 0.0], atol=1e-15)
    v = np.array([2.0 + 0j, 2.0 + 0j]) * np.exp(1j * 0.4)
    out = po.mmp1_constraint(v)
    assert np.allclose(out, v, atol=1e-14)
