"""Seeded random configurations against the FP64 oracle (needs a GPU): per
seed, one plan holding twelve items of random variant, channel count, length,
sampling rate, energyError, iteration limit and dictionary options, each
compared with its own oracle run (indices up to the first near-tie,
amplitudes, phases and energies); and the learned proposal at delta = 0 on
random single-channel items, which must pass the same parity.  The draws are
fixed by the seed, so the cases are the same on every run."""
import numpy as np
import pytest

import empi_amd as E
from empi_amd import synth
from test_gpu_parity import _announce, _check_item, _oracle, _run_plan

pytestmark = pytest.mark.gpu

VARIANTS = [E.SMP, E.MMP1, E.MMP2, E.MMP3]


def _cases(seed=7100, count=12):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        variant = VARIANTS[i % 4]
        C = 1 if variant == E.SMP and i % 8 == 0 else int(rng.integers(2, 5))
        N = int(rng.integers(300, 2500))
        fs = float(rng.choice([64.0, 100.0, 128.0, 250.0, 256.0]))
        eps = float(rng.choice([0.005, 0.01, 0.02, 0.05]))
        iters = int(rng.integers(3, 16))
        percent = float(rng.choice([90.0, 99.0, 99.9]))
        kw = {}
        if i % 3 == 1:
            kw["freq_max"] = float(fs * rng.uniform(0.15, 0.45))
        if i % 5 == 2:
            kw["scale_max"] = float(N / fs * rng.uniform(0.2, 0.8))
        x = synth.shared_multichannel(C=C, N=N, fs=fs, atoms=int(rng.integers(1, 5)), noise_sigma=0.2,
                                      seed=int(rng.integers(1 << 30)))
        out.append((variant, C, N, fs, eps, iters, percent, kw, x))
    return out


@pytest.mark.parametrize("seed", [7100, 7200, 7300])
def test_random_configurations_one_plan(seed):
    cases = _cases(seed)
    _announce(f"{len(cases)} random items (seed {seed}) in one plan, every variant")
    specs = [E.ItemSpec(C, N, fs, v, it, pc, eps, scale_min=0.0,
                        scale_max=kw.get("scale_max", np.inf), freq_max=kw.get("freq_max", np.inf))
             for (v, C, N, fs, eps, it, pc, kw, x) in cases]
    plan = _run_plan(specs, [c[-1] for c in cases])
    for i, (v, C, N, fs, eps, it, pc, kw, x) in enumerate(cases):
        where = f"item {i}: variant {v}, {C} ch x {N} @{fs:g} Hz, eps {eps}, {it} it, {pc}%, {kw}"
        print(where, flush=True)
        res = plan.result(i)
        assert res.status == 0, where
        ora = _oracle(x, fs, eps, v, it, pc, **kw)
        _check_item(res, ora, v, where)


@pytest.mark.parametrize("seed", [7400, 7500])
def test_random_configurations_proposal_exact(seed):
    """The learned proposal at delta = 0 (every iteration certified exact) on
    random single-channel items: the atoms must pass the same oracle parity
    as the plain method."""
    from empi_amd import proposal
    rng = np.random.default_rng(seed)
    _announce(f"proposal delta=0 on 10 random single-channel items (seed {seed})")
    cases = []
    for _ in range(10):
        N = int(rng.integers(600, 4000))
        fs = float(rng.choice([128.0, 250.0, 256.0]))
        eps = float(rng.choice([0.01, 0.02]))
        iters = int(rng.integers(5, 20))
        x = synth.pink_with_atoms(int(rng.integers(1 << 30)), N, fs, int(rng.integers(1, 7)))
        cases.append((N, fs, eps, iters, x))
    specs = [E.ItemSpec(1, N, fs, E.SMP, it, 99.0, eps) for (N, fs, eps, it, x) in cases]
    plan = E.Plan(specs)
    plan.set_proposal(E.PROPOSAL_ON, proposal.load(), top_k=8, delta=0.0)
    plan.run(np.concatenate([c[-1] for c in cases]).astype(np.float32))
    for i, (N, fs, eps, it, x) in enumerate(cases):
        where = f"item {i}: {N} @{fs:g} Hz, eps {eps}, {it} it"
        res = plan.result(i)
        assert res.status == 0, where
        ora = _oracle(x[None, :], fs, eps, E.SMP, it, 99.0)
        _check_item(res, ora, E.SMP, where)
    print("stats", plan.proposal_stats(), flush=True)
