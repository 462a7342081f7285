"""Indexed max-heap of the oracle against the reference's heap.hpp compiled
from its own source (oracle/_ref/heap_driver, built by `make -C oracle ref`
where the reference tree is present) and against sorting."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def test_heap_peek_is_sorted_order():
    """The reference's own property test (tests/test-heap.cpp:13-48), restated."""
    rng = np.random.default_rng(1)
    n = 1000
    values = rng.random(n)
    h = po.Heap(values)
    for _ in range(6):
        order = np.argsort(-values, kind="stable")
        for key in order:
            k, v = h.peek()
            assert v == values[key]
            assert values[k] == v
            h.update(k, -1.0)
            values[k] = -1.0
        assert h.peek()[1] == -1.0
        keys = rng.permutation(n)
        for key in keys:
            values[key] = rng.random()
            h.update(key, values[key])


def _ops(seed, n, steps):
    rng = np.random.default_rng(seed)
    vals = rng.integers(0, 6, n).astype(float)     # many exact ties
    ops = []
    for _ in range(steps):
        if rng.random() < 0.3:
            ops.append(("p",))
        else:
            ops.append(("u", int(rng.integers(0, n)), float(rng.integers(-2, 8))))
    ops.append(("p",))
    return vals, ops


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "heap_driver")), reason="reference heap driver not built")
@pytest.mark.parametrize("seed,n", [(0, 2), (1, 3), (2, 17), (3, 100), (4, 1000)])
def test_heap_matches_reference_on_ties(seed, n):
    vals, ops = _ops(seed, n, 3000)
    text = f"{n} " + " ".join(repr(float(v)) for v in vals) + "\n"
    text += "\n".join("p" if o[0] == "p" else f"u {o[1]} {o[2]!r}" for o in ops) + "\n"
    out = subprocess.run([os.path.join(REF, "heap_driver")], input=text, capture_output=True, text=True,
                         check=True).stdout.split("\n")
    h = po.Heap(vals)
    got = []
    for o in ops:
        if o[0] == "p":
            k, v = h.peek()
            got.append(f"{k} {v:.17g}")
        else:
            h.update(o[1], o[2])
    assert got == [line for line in out if line]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "test-heap")), reason="reference heap test not built")
def test_reference_heap_test_passes():
    assert subprocess.run([os.path.join(REF, "test-heap")]).returncode == 0
