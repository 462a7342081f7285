"""Parity rules shared by the GPU tests.

Atom indices (scale, bin, time) must match the FP64 oracle exactly at every
iteration before the first one whose oracle top-two products differ by
1e-4 relative or less; amplitudes and phases within 1e-4 relative and
residual energies within 1e-5 relative over the same iterations.
"""
import numpy as np

GAP = 1e-4
AMP_RTOL = 1e-4
PHASE_ATOL = 1e-4 * np.pi
ENERGY_RTOL = 1e-5


def compared_iterations(gaps_row, count):
    """Iterations [0, n) before the first near-tie of the oracle."""
    n = 0
    while n < count and not (gaps_row[n] <= GAP):
        n += 1
    return n


def wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


def compare_rows(gpu_params, gpu_index, ora_params, ora_index, n):
    """Returns a dict of the worst errors over the first n atoms of one row."""
    out = {"n": n, "index_mismatch": 0, "amp_rel": 0.0, "phase_abs": 0.0, "mod_rel": 0.0}
    for i in range(n):
        if tuple(gpu_index[i]) != tuple(ora_index[i]):
            out["index_mismatch"] += 1
            continue
        g, o = gpu_params[i], ora_params[i]
        out["amp_rel"] = max(out["amp_rel"], abs(g[1] - o[1]) / max(abs(o[1]), 1e-300))
        out["mod_rel"] = max(out["mod_rel"], abs(g[0] - o[0]) / max(abs(o[0]), 1e-300))
        out["phase_abs"] = max(out["phase_abs"], abs(wrap(g[5] - o[5])))
        for j in (2, 3, 4):
            assert abs(g[j] - o[j]) <= 1e-9 * max(1.0, abs(o[j])), (i, j, g[j], o[j])
    return out


def assert_parity(err, where=""):
    assert err["index_mismatch"] == 0, (where, err)
    assert err["amp_rel"] <= AMP_RTOL, (where, err)
    assert err["mod_rel"] <= AMP_RTOL, (where, err)
    assert err["phase_abs"] <= PHASE_ATOL, (where, err)


def energy_rel(gpu_e, ora_e, n):
    m = min(n + 1, len(ora_e))
    g, o = np.asarray(gpu_e[:m]), np.asarray(ora_e[:m])
    ok = ~np.isnan(o)
    if not ok.any():
        return 0.0
    return float(np.max(np.abs(g[ok] - o[ok]) / np.abs(o[ok])))
