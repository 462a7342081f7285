"""Windows longer than 8192 samples (needs a GPU): the packed input of such a
window spans several 4096-point sections, which fold onto one sub-transform
(input m1 + 4096 f enters sub-transform k2 with e^{-2 pi i k2 f / P}).  The
largest scales of a 16,384-sample signal fold twice, of a 32,768-sample
signal four times; every class runs in the same plan.  Parity against the
FP64 oracle, as for the short signals."""
import numpy as np

import empi_amd as E
import pyoracle as po
from empi_amd import synth
from test_gpu_parity import _announce, _check_item, _oracle, _run_plan

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N, iters", [(16384, 8), (32768, 4)])
def test_folded_windows_match_oracle(N, iters):
    _announce(f"long windows: pink noise + 6 atoms, {N} samples @256 Hz, {iters} iterations")
    x = synth.pink_with_atoms(4300 + N // 16384, N, 256.0, 6)
    spec = E.ItemSpec(1, N, 256.0, E.SMP, iters, 99.0, 0.01)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(x[None, :], 256.0, 0.01, po.SMP, iters, 99.0)
    _check_item(res, ora, E.SMP, f"{N} samples")
