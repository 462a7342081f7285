"""empi_decompose_noalloc (needs a GPU): the allocation-free entry point
against the plan path (empi_plan_create + empi_plan_run +
empi_plan_item_result) on the same inputs, through the C ABI, on two calls
in a row (the second with a smaller input, so it runs over the first call's
scratch contents), with device and host inputs, plus the size limit."""
import numpy as np
import pytest
import torch

import empi_amd as E
from empi_amd import synth

pytestmark = pytest.mark.gpu


def _plan_result(spec, x):
    plan = E.Plan([spec])
    plan.run(np.asarray(x, np.float32).ravel())
    res = plan.result(0)
    plan.close()
    return res


def _same(a, b, where):
    assert a.status == b.status, where
    assert np.array_equal(a.counts, b.counts), where
    assert np.array_equal(a.index, b.index), where
    assert np.array_equal(a.params, b.params), where
    assert np.array_equal(a.energies, b.energies, equal_nan=True), where


def test_noalloc_matches_plan_two_calls():
    print("\n[gpu case] empi_decompose_noalloc: MMP3 4 x 4096 then SMP 1 x 1280 vs the plan path", flush=True)
    big = E.ItemSpec(4, 4096, 256.0, E.MMP3, 20, 99.5, 0.01)
    xb = synth.shared_multichannel(C=4, N=4096, fs=256.0, atoms=4, noise_sigma=0.1, seed=71)
    small = E.ItemSpec(1, 1280, 128.0, E.SMP, 30, 99.0, 0.01)
    xs = synth.single_channel()[None, :]
    want_b = _plan_result(big, xb)
    want_s = _plan_result(small, xs)
    db = torch.from_numpy(np.ascontiguousarray(xb, np.float32).ravel()).to("cuda:0")
    ds = torch.from_numpy(np.ascontiguousarray(xs, np.float32).ravel()).to("cuda:0")
    torch.cuda.synchronize()
    got_b = E.decompose_noalloc(big, device_ptr=db.data_ptr())
    got_s = E.decompose_noalloc(small, device_ptr=ds.data_ptr())
    _same(got_b, want_b, "first call (MMP3 4 x 4096)")
    _same(got_s, want_s, "second call (SMP 1280)")
    assert int(got_b.counts[0]) > 1 and int(got_s.counts[0]) > 1
    # host input, and the larger item again after the smaller one
    _same(E.decompose_noalloc(small, signal=xs), want_s, "host input")
    _same(E.decompose_noalloc(big, device_ptr=db.data_ptr()), want_b, "third call (MMP3 again)")


def test_noalloc_uneven_smp_channels():
    print("\n[gpu case] empi_decompose_noalloc: SMP 3 x 2000 vs the plan path", flush=True)
    spec = E.ItemSpec(3, 2000, 256.0, E.SMP, 25, 99.0, 0.01)
    x = np.stack([synth.pink_with_atoms(880 + c, 2000, 256.0, 5) for c in range(3)])
    _same(E.decompose_noalloc(spec, signal=x), _plan_result(spec, x), "SMP 3 x 2000")


def test_noalloc_rejects_items_beyond_the_scratch():
    """64 channels x 10240 samples of MMP3 need far more than 64 MiB of cell
    bests: unsupportedSize, decided on the host before any upload."""
    print("\n[gpu case] empi_decompose_noalloc: oversize item rejected", flush=True)
    spec = E.ItemSpec(64, 10240, 512.0, E.SMP, 10, 99.0, 0.005)
    with pytest.raises(E.EmpiError) as ei:
        E.decompose_noalloc(spec, signal=np.zeros((64, 10240), np.float32))
    assert ei.value.status == 8
