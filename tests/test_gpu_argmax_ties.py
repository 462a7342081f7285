"""Exact ties in the device argmax (needs a GPU).

The engine orders map entries by FP32 product and breaks exact ties by the
smallest scale, then bin, then time index, through three keys: per cell the
packed (product, ~bin) of its bins, per tile the (product, ~(bin, cell)) of
its cells, and the selection's (product, scale, bin, time) over tiles
(device_math.hpp: pack_best, tile_key, sel_of_tile, sel_better; the column
kernels and k_select use the same functions).  The reference's heap
(inc/heap.hpp) leaves the order of exact ties to its update history, and its
own test (tests/test-heap.cpp:13-48) checks distinct values only.

empi_debug_argmax runs those keys on the device over a planted map, laid out
as the column kernels lay out a scale: time index t, cells g < G = P/2 per
time index (g = 0 also holds sub-transform P/2), bins k = P k1 + k2 of the
cell's sub-transforms k2.  Products drawn from a few small integers tie at
every level; the result must be the brute-force maximum with that tie order.
Negative and NaN products never win; a map with no positive product has best
0.
"""
import numpy as np
import pytest

import empi_amd as E
from test_gpu_parity import _announce

pytestmark = pytest.mark.gpu


def _bins(P, Q):
    """[G][2Q] bin indices of one column's cells (P >= 2)."""
    G = P // 2
    out = np.zeros((G, 2 * Q), np.int64)
    for g in range(G):
        k2s = (0, P // 2) if g == 0 else (g, P - g)
        out[g] = np.concatenate([P * np.arange(Q) + k2 for k2 in k2s])
    return out


def _brute(maps):
    """((scale, bin, time), best) by brute force: product desc, then scale,
    bin, time asc; negative and NaN products count as 0."""
    best, arg = 0.0, None
    cands = []
    for l, (p, bins, _) in enumerate(maps):
        q = np.where(np.isnan(p) | (p < 0), 0.0, p).astype(np.float32)
        m = float(q.max())
        if m > best:
            best, cands = m, []
        if m == best and m > 0:
            t, g, b = np.nonzero(q == m)
            cands += [(l, int(bins[gg, bb]), int(tt)) for tt, gg, bb in zip(t, g, b)]
    if best > 0:
        arg = min(cands)
    return arg, best


def _check(maps, where):
    got, best = E.debug_argmax(maps)
    want, wbest = _brute(maps)
    assert best == wbest, (where, best, wbest)
    if wbest > 0:
        assert got == want, (where, got, want)
    return want


def _random_maps(rng, levels, shapes):
    maps = []
    for P, Q, T, nc in shapes:
        bins = _bins(P, Q)
        p = rng.integers(0, levels, size=(T, P // 2, 2 * Q)).astype(np.float32)
        maps.append((p, bins, nc))
    return maps


def test_random_planted_ties():
    """Products from {0, 1, 2} (ties everywhere) and {0 .. 9}: 300 maps of
    three scales with cells per tile from 1 to 256."""
    _announce("argmax ties: 300 planted maps of 3 scales vs brute force")
    rng = np.random.default_rng(7100)
    shapes = [[(2, 8, 40, 256), (4, 16, 33, 16), (8, 32, 9, 4)],
              [(2, 4, 70, 32), (2, 64, 12, 1), (16, 8, 20, 8)],
              [(4, 4, 100, 256), (8, 8, 17, 64), (2, 128, 5, 2)]]
    tied = 0
    for trial in range(300):
        maps = _random_maps(rng, 3 if trial % 2 == 0 else 10, shapes[trial % 3])
        _check(maps, f"trial {trial}")
        top = max(float(m[0].max()) for m in maps)
        tied += sum(int((m[0] == top).sum()) for m in maps) > 1
    assert tied >= 290, tied   # nearly every trial has an exact tie at the top


def test_ties_at_each_level():
    """One maximum value planted at two places: two bins of one cell, two
    cells of one tile (same bin, different time), two tiles, two scales;
    every case has the smaller key second in memory order too."""
    _announce("argmax ties: planted pairs at bin, cell, tile and scale level")
    rng = np.random.default_rng(7200)
    base = _random_maps(rng, 3, [(4, 16, 24, 8), (2, 32, 30, 16)])
    top = np.float32(5.0)

    def plant(places):
        maps = [(m[0].copy(), m[1], m[2]) for m in base]
        for l, t, g, b in places:
            maps[l][0][t, g, b] = top
        return maps

    b0, b1 = base[0][1], base[1][1]   # [G][2Q] bins of scales 0 (P = 4) and 1 (P = 2)
    # two bins of one cell (time 3, g = 1): the smaller bin wins
    assert _check(plant([(0, 3, 1, 20), (0, 3, 1, 2)]), "one cell") == (0, int(min(b0[1, 20], b0[1, 2])), 3)
    # two cells of one tile (8 cells per tile, G = 2: times 4..7), same bin: time 5 beats time 6
    assert _check(plant([(0, 6, 0, 5), (0, 5, 0, 5)]), "one tile") == (0, int(b0[0, 5]), 5)
    # two tiles, same bin: time 2 (tile 0) beats time 21 (tile 5)
    assert _check(plant([(0, 21, 1, 7), (0, 2, 1, 7)]), "two tiles") == (0, int(b0[1, 7]), 2)
    # two scales: scale 0 wins over scale 1 even at a larger bin and time
    assert _check(plant([(1, 0, 0, 0), (0, 23, 1, 31)]), "two scales") == (0, int(b0[1, 31]), 23)
    # equal value: the smaller bin at the later time wins
    assert _check(plant([(0, 1, 0, 30), (0, 20, 0, 1)]), "bin before time") == (0, int(b0[0, 1]), 20)
    assert b0[0, 1] < b0[0, 30] and b1.shape == (1, 64)


def test_non_positive_products_never_win():
    _announce("argmax ties: negative and NaN products, empty map")
    rng = np.random.default_rng(7300)
    maps = _random_maps(rng, 3, [(4, 8, 10, 4), (2, 16, 12, 8)])
    for p, _, _ in maps:
        p[:] = -1.0
        p[::2, :, ::3] = np.nan
    got, best = E.debug_argmax(maps)
    assert best == 0.0
    maps[1][0][7, 0, 9] = 0.5
    maps[0][0][3, 1, 2] = np.nan
    assert _check(maps, "one positive") == (1, int(maps[1][1][0, 9]), 7)


def test_widest_keys():
    """Bins up to 2^24 - 1 (24 bits of the tile key) and 256 cells per tile
    (8 bits): the largest bin of the last cell of a full tile."""
    _announce("argmax ties: bins near 2^24, 256 cells per tile")
    P, Q = 2, 8
    bins = _bins(P, Q) + ((1 << 24) - 16)   # bins 2^24 - 16 .. 2^24 - 1
    assert bins.max() == (1 << 24) - 1
    p = np.zeros((600, 1, 2 * Q), np.float32)
    p[511, 0, 15] = 3.0   # bin 2^24 - 1, cell 255 of tile 1
    p[255, 0, 15] = 3.0   # bin 2^24 - 1, cell 255 of tile 0
    assert _check([(p, bins, 256)], "largest bin, last cell") == (0, (1 << 24) - 1, 255)
    p[300, 0, 14] = 3.0   # bin 2^24 - 3 at a later time wins
    assert _check([(p, bins, 256)], "smaller bin") == (0, (1 << 24) - 3, 300)
