"""Long windows (needs a GPU).  A single-channel unit's window longer than
8192 samples spans several 4096-point sections of packed input, which fold
onto one sub-transform (input m1 + 4096 f enters sub-transform k2 with
e^{-2 pi i k2 f / P}); multichannel units fold beyond 16,384 samples.  The
largest scales of a 16,384-sample signal fold twice, of a 32,768-sample
signal four times, of a 300,000-sample signal 37 times; every class runs in
the same plan.  Parity against the FP64 oracle, as for the short signals."""
import pytest

import empi_amd as E
import pyoracle as po
from empi_amd import synth
from test_gpu_parity import _announce, _check_item, _oracle, _run_plan

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


@pytest.mark.parametrize("variant", [E.MMP3, E.MMP1])
def test_folded_windows_multichannel(variant):
    """Two channels of 20,000 samples: multichannel units run windows up to
    16,384 samples unfolded (8192-point sub-transforms) and fold the longer
    ones three times onto 4096 points, in the sum and MMP1 modes."""
    _announce(f"long windows, 2 channels x 20000 samples @256 Hz, variant {variant}, 3 iterations")
    x = synth.shared_multichannel(C=2, N=20000, fs=256.0, atoms=4, noise_sigma=0.1, seed=4400)
    spec = E.ItemSpec(2, 20000, 256.0, variant, 3, 99.0, 0.01)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(x, 256.0, 0.01, variant, 3, 99.0)
    _check_item(res, ora, variant, f"2 x 20000 samples, variant {variant}")


def test_nine_folds():
    """70,000 samples: the largest windows span 35,000 packed points, nine
    4096-point sections."""
    _announce("long windows: 70000 samples @256 Hz, 2 iterations (nine folds)")
    x = synth.pink_with_atoms(4350, 70000, 256.0, 4)
    spec = E.ItemSpec(1, 70000, 256.0, E.SMP, 2, 99.0, 0.01)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(x[None, :], 256.0, 0.01, po.SMP, 2, 99.0)
    _check_item(res, ora, E.SMP, "70000 samples")


def test_thirty_seven_folds():
    """300,000 samples at 64 Hz (energyError 0.1 keeps the oracle's dictionary
    small): the largest windows span 150,000 packed points, 37 sections,
    past the 16-section restarts of the section twiddles."""
    _announce("long windows: 300000 samples @64 Hz, eps 0.1, 3 iterations (37 folds)")
    x = synth.pink_with_atoms(4360, 300000, 64.0, 4)
    spec = E.ItemSpec(1, 300000, 64.0, E.SMP, 3, 99.0, 0.1)
    plan = _run_plan([spec], [x])
    res = plan.result(0)
    assert res.status == 0
    ora = _oracle(x[None, :], 64.0, 0.1, po.SMP, 3, 99.0)
    _check_item(res, ora, E.SMP, "300000 samples", min_compared=3)
