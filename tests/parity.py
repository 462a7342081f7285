"""Parity rules shared by the GPU tests: the engine against the FP64 oracle
on the same input.

* Atom indices (scale, bin, time) match at every iteration wherever the
  oracle's top-two products differ by more than 1e-4 relative.  Where they
  differ by less, the two may take different branches of the near-tie, so the
  oracle runs *following* the engine's choices (eo_decompose_follow): it
  takes the engine's atom when that atom's product is within 1e-4 of its own
  best and keeps its own otherwise.  Every iteration of every decomposition is
  therefore compared, and a choice more than 1e-4 below the best is an index
  mismatch.
* Amplitudes and moduli agree within 1e-4 relative, phases within 1e-4 pi,
  position, scale and frequency within 1e-9.
* Residual energies agree within 1e-5 relative at every iteration, and the
  atom counts are equal (same stop).
"""
import numpy as np

import pyoracle as po

GAP = 1e-4
AMP_RTOL = 1e-4
PHASE_ATOL = 1e-4 * np.pi
ENERGY_RTOL = 1e-5


def follow_rows(index, counts, variant, iters):
    """The engine's (k, f, t) per oracle row, [R][iters][3] (-1 past the
    count): one row per channel for SMP, channel 0 otherwise."""
    index = np.asarray(index)
    rows = index.shape[0] if variant == po.SMP else 1
    out = np.full((rows, iters, 3), -1, np.int32)
    for r in range(rows):
        n = min(int(counts[r]), iters, index.shape[1])
        out[r, :n] = index[r, :n]
    return out


def oracle_following(d, x, variant, iters, percent, index, counts, threads=0, iter_limit=None):
    """The oracle on x, following the engine's atoms through near-ties."""
    return po.decompose(d, np.asarray(x, np.float64), variant, iters, percent, threads=threads,
                        iter_limit=iter_limit, follow=follow_rows(index, counts, variant, iters), follow_tol=GAP)


def wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


def compare_rows(gpu_params, gpu_index, gpu_count, ora_params, ora_index, ora_count, followed=None,
                 limited=False):
    """Worst errors over one channel's atoms.  limited: the oracle stopped at an
    iteration limit, so only its atoms are compared and the counts are not."""
    n = int(ora_count) if limited else min(int(gpu_count), int(ora_count))
    out = {"n": n, "count_gpu": int(gpu_count), "count_oracle": int(ora_count), "limited": limited,
           "index_mismatch": 0, "first_mismatch": None, "followed": 0,
           "amp_rel": 0.0, "phase_abs": 0.0, "mod_rel": 0.0}
    if followed is not None:
        out["followed"] = int(np.sum(followed[:n]))
    for i in range(n):
        if tuple(gpu_index[i]) != tuple(ora_index[i]):
            out["index_mismatch"] += 1
            if out["first_mismatch"] is None:
                out["first_mismatch"] = (i, tuple(int(v) for v in gpu_index[i]), tuple(int(v) for v in ora_index[i]))
            continue
        g, o = gpu_params[i], ora_params[i]
        out["amp_rel"] = max(out["amp_rel"], abs(g[1] - o[1]) / max(abs(o[1]), 1e-300))
        out["mod_rel"] = max(out["mod_rel"], abs(g[0] - o[0]) / max(abs(o[0]), 1e-300))
        out["phase_abs"] = max(out["phase_abs"], abs(wrap(g[5] - o[5])))
        for j in (2, 3, 4):
            assert abs(g[j] - o[j]) <= 1e-9 * max(1.0, abs(o[j])), (i, j, g[j], o[j])
    return out


def assert_parity(err, where="", min_compared=1):
    assert err["index_mismatch"] == 0, (where, err)
    if not err["limited"]:
        assert err["count_gpu"] == err["count_oracle"], (where, err)
    assert err["n"] >= min_compared, (where, err)
    assert err["amp_rel"] <= AMP_RTOL, (where, err)
    assert err["mod_rel"] <= AMP_RTOL, (where, err)
    assert err["phase_abs"] <= PHASE_ATOL, (where, err)


def energy_rel(gpu_e, ora_e, n):
    """Worst relative difference of the residual energies after 0..n atoms."""
    m = min(n + 1, len(ora_e))
    g, o = np.asarray(gpu_e[:m]), np.asarray(ora_e[:m])
    ok = ~np.isnan(o)
    assert np.array_equal(ok, ~np.isnan(g)), (g, o)
    if not ok.any():
        return 0.0
    return float(np.max(np.abs(g[ok] - o[ok]) / np.abs(o[ok])))


def energy_drop_identity(fs, N, params, energies, count):
    """Worst relative error of the per-iteration identity
    fs (E_i - E_{i+1}) = sum_c (2 modulus_ci^2 - ||w_ci||^2): the chosen map
    value gives <r_c, w_c> = modulus_c^2 exactly, whatever the clipping at the
    edges.  params [C][I][6] of one row; iterations with both energies known."""
    params = np.asarray(params, np.float64)
    if params.ndim == 2:
        params = params[None]
    C = params.shape[0]
    e = np.asarray(energies[:count + 1], np.float64)
    m = int(np.sum(~np.isnan(e))) - 1      # subtractions recorded
    if m <= 0:
        return 0.0
    dE = fs * (e[:m] - e[1:m + 1])
    w2 = np.stack([po.waveform_norm2(fs, N, params[c, :m]) for c in range(C)])
    pred = np.sum(2.0 * params[:, :m, 0] ** 2 - w2, axis=0)
    return float(np.max(np.abs(dE - pred) / np.maximum(np.abs(pred), 1e-300)))
