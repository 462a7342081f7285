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
    bad[classes] = 1     # multi-channel classes take grid variants only
    with pytest.raises(E.EmpiError):
        plan.set_tuning(bad)
