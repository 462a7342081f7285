"""Column-kernel autotuner (needs a GPU): every variant computes the same
products, so the atoms of a run do not depend on the tuning table."""
import numpy as np
import pytest

import empi_amd as E
from empi_amd import synth
from test_gpu_parity import _announce

pytestmark = pytest.mark.gpu


def _atoms(plan, n):
    return [plan.result(i).params.copy() for i in range(n)], [plan.result(i).index.copy() for i in range(n)]


def test_autotune_keeps_results():
    _announce("autotune: 3 uneven signals, 20 iterations, every variant")
    xs = synth.uneven_batch(items=3, nmax=6000)
    specs = [E.ItemSpec(1, len(x), 256.0, E.SMP, 20, 99.0, 0.01) for x in xs]
    flat = np.concatenate(xs).astype(np.float32)
    plan = E.Plan(specs)
    plan.run(flat)
    p0, i0 = _atoms(plan, len(xs))
    table = plan.autotune(flat, repeats=1)
    print("tuning", table.tolist(), flush=True)
    classes = E.lib().empi_plan_kernel_classes()
    assert table.size == 3 * classes and np.all((table >= 0) & (table < 6))
    plan.run(flat)
    p1, i1 = _atoms(plan, len(xs))
    for a, b in zip(i0, i1):
        assert np.array_equal(a, b)
    for a, b in zip(p0, p1):
        assert np.array_equal(a, b)
    # every single-channel variant, forced
    for v in range(6):
        t = table.copy()
        t[:classes] = v
        plan.set_tuning(t)
        plan.run(flat)
        p2, i2 = _atoms(plan, len(xs))
        for a, b in zip(i0, i2):
            assert np.array_equal(a, b), v
    bad = table.copy()
    bad[classes] = 6     # six variants per class
    with pytest.raises(E.EmpiError):
        plan.set_tuning(bad)


def test_multichannel_variants_keep_results():
    """Every wave and grid variant of the sum (MMP3) and MMP1 classes gives the
    same atoms."""
    _announce("autotune: MMP3 and MMP1 items, 4 ch x 2048, 12 iterations, every multichannel variant")
    specs = [E.ItemSpec(4, 2048, 256.0, E.MMP3, 12, 99.9, 0.01), E.ItemSpec(4, 2048, 256.0, E.MMP1, 12, 99.9, 0.01)]
    xs = [synth.shared_multichannel(C=4, N=2048, fs=256.0, atoms=4, noise_sigma=0.1, seed=40 + i) for i in range(2)]
    flat = np.concatenate([x.ravel() for x in xs]).astype(np.float32)
    plan = E.Plan(specs)
    plan.run(flat)
    p0, i0 = _atoms(plan, 2)
    classes = E.lib().empi_plan_kernel_classes()
    table = plan.tuning()
    for v in range(6):
        t = table.copy()
        t[classes:] = v
        plan.set_tuning(t)
        plan.run(flat)
        p2, i2 = _atoms(plan, 2)
        for a, b in zip(i0, i2):
            assert np.array_equal(a, b), v
        for a, b in zip(p0, p2):
            assert np.array_equal(a, b), v
    tuned = plan.autotune(flat, repeats=1)
    print("tuning", tuned.tolist(), flush=True)
    plan.run(flat)
    p3, _ = _atoms(plan, 2)
    for a, b in zip(p0, p3):
        assert np.array_equal(a, b)
