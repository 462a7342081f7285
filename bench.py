"""Benchmark: batched SMP decomposition on one MI355X.

Workload: 64 channels x 20 epochs x 5120 float32 samples at 256 Hz, each
channel-epoch unit-variance pink noise + 8 random Gabor atoms; optimal Gabor
dictionary with energyError 0.01, 100 iterations, energyPercent 99; FP32
maps, FP64 energy bookkeeping.  A step decomposes the whole batch (full map
build + MP loop) with the input already resident in HBM.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--channels", type=int, default=64)
    ap.add_argument("--iterations", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iterations", type=int, default=20,
                    help="iterations of the CPU-baseline sample (one channel-epoch)")
    return ap.parse_args()


def cpu_baseline(x0, fs, iters_total, cpu_iters):
    """FP64 oracle (OpenMP over scales, own FFT) on one channel-epoch,
    first cpu_iters iterations: atoms/s on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = min(16, os.cpu_count() or 1)
    d = po.Dictionary(fs, len(x0), 0.01)
    t0 = time.perf_counter()
    r = po.decompose(d, x0[None, :].astype(np.float64), po.SMP, iters_total, 99.0,
                     threads=threads, iter_limit=cpu_iters)
    dt = time.perf_counter() - t0
    atoms = int(r.counts.sum())
    return {"value": atoms / dt, "unit": "atoms/s", "cores": threads, "kind": "port",
            "sample": f"1 of the channel-epochs (5120 samples), full map build + first {atoms} "
                      f"iterations, {dt:.1f} s, FP64 oracle with its own radix-2 FFT"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch
    import empi_amd as E
    from empi_amd import synth

    fs, N = 256.0, 5120
    items = args.epochs
    C = args.channels
    I = args.iterations
    sig = synth.batched_channels(items=items * C, N=N, fs=fs, atoms=8)   # [items*C][N]
    specs = [E.ItemSpec(C, N, fs, E.SMP, I, 99.0, 0.01) for _ in range(items)]
    plan = E.Plan(specs)
    dev = torch.from_numpy(np.ascontiguousarray(sig, dtype=np.float32)).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()

    def step():
        plan.run(device_ptr=ptr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    atoms = 0
    explained = []
    for i in range(items):
        r = plan.result(i)
        if r.status != 0:
            raise RuntimeError(f"item {i}: {E.STATUS_NAMES.get(r.status)}")
        atoms += int(r.counts.sum())
        for c in range(C):
            e = r.energies[c]
            last = e[~np.isnan(e)][-1]
            explained.append(1.0 - last / e[0])
    value = atoms * args.steps / dt

    # per-class kernel time with HIP events on the launch stream (one profiled step)
    plan.set_profiling(True)
    step()
    plan.set_profiling(False)
    stats = plan.kernel_stats()
    dom = max(stats, key=lambda s: s["total_ms"])
    bytes_per_entry = 8 * 1 + 4          # one FP32 complex map value per channel + one FP32 product
    avg_ms = dom["total_ms"] / dom["launches"]
    bytes_per_launch = dom["entries"] * bytes_per_entry / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": f"k_columns<{dom['sub_length']}>", "avg_launch_ms": avg_ms,
                "entries_per_launch": dom["entries"] / dom["launches"]}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sig[0].astype(np.float64), fs, I, args.cpu_iterations)

    out = {
        "metric": "atoms/sec to 99% energy on 1 MI355X; proposal vs plain at stated energy error",
        "value": value, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 maps / fp64 energy", "data": "synthetic",
        "config": {"workload": f"SMP batched plain method: {C} channels x {items} epochs x {N} "
                               f"float32 samples at 256 Hz, pink noise + 8 Gabor atoms, "
                               f"energyError 0.01, {I} iterations, energyPercent 99",
                   "method": "plain", "channel_epochs": items * C, "iterations_max": I,
                   "atoms_per_step": atoms, "mean_explained_energy": float(np.mean(explained))},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": stats,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
