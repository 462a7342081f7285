"""Benchmark: Matching Pursuit decomposition on one MI355X.

Default workload (--workload smp-batch): 64 channels x 20 epochs x 5120
float32 samples at 256 Hz, each channel-epoch unit-variance pink noise + 8
random Gabor atoms; optimal Gabor dictionary with energyError 0.01, 100
iterations, energyPercent 99; FP32 maps, FP64 energy bookkeeping.  A step
decomposes the whole batch (full map build + MP loop) with the input already
resident in HBM.  Other workloads:
  single   1280 samples at 128 Hz, 5 atoms + noise 0.1, 50 iterations (SMP)
  mmp3     32 channels x 10240 samples at 512 Hz, 12 shared atoms, MMP3,
           energyError 0.005, 200 iterations
  uneven   512 single-channel signals of 1024..16384 samples at 256 Hz, pink
           noise + Poisson(6) atoms, 99% / at most 100 iterations

With the default workload the line also carries "proposal": the learned
proposal against the plain method on the uneven batch (atoms/s of both,
entries evaluated, explained-energy deficit per signal).

Prints one JSON line on stdout (progress goes to stderr).  --gpus above 1 is
refused: this build runs on one GPU.
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
METRIC = "atoms/sec to 99% energy on 1 MI355X; proposal vs plain at stated energy error"


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="smp-batch",
                    choices=["smp-batch", "single", "mmp3", "uneven", "proposal", "copy-calibration"])
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--channels", type=int, default=64)
    ap.add_argument("--iterations", type=int, default=None)
    ap.add_argument("--items", type=int, default=512, help="signals of the uneven workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iterations", type=int, default=20,
                    help="iterations of the CPU-baseline sample (smp-batch)")
    ap.add_argument("--no-proposal", action="store_true",
                    help="skip the learned-proposal comparison on the uneven batch")
    ap.add_argument("--proposal-items", type=int, default=512)
    ap.add_argument("--proposal-delta", type=float, default=1e-4)
    ap.add_argument("--tuning", default=None,
                    help="comma-separated column-kernel tuning table instead of autotuning")
    return ap.parse_args()


def proposal_leg(args, torch):
    """Learned proposal vs the plain method on the uneven batch (512 signals
    of 1024..16384 samples at 256 Hz, pink noise + Poisson(6) atoms, 99% /
    at most 100 iterations): atoms/s of both (one untimed run each first),
    map entries evaluated, and the proposal's explained-energy deficit
    against the plain method's, per signal (max, mean) and of the whole batch
    (explained energy summed over the signals)."""
    import empi_amd as E
    from empi_amd import proposal, synth
    xs = synth.uneven_batch(items=args.proposal_items)
    specs = [E.ItemSpec(1, len(x), 256.0, E.SMP, 100, 99.0, 0.01) for x in xs]
    dev = torch.from_numpy(np.concatenate([x.astype(np.float32) for x in xs])).to("cuda:0")
    net = proposal.load()
    out = {}
    tuning = None
    for name in ("plain", "proposal"):
        plan = E.Plan(specs)
        if tuning is None:
            tuning = plan.autotune(device_ptr=dev.data_ptr(), repeats=2)
        else:
            plan.set_tuning(tuning)
        if name == "proposal":
            plan.set_proposal(E.PROPOSAL_ON, net, top_k=8, delta=args.proposal_delta)
        plan.run(device_ptr=dev.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run(device_ptr=dev.data_ptr())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        atoms = sum(int(plan.result(i).counts.sum()) for i in range(len(specs)))
        expl, removed = [], []
        for i in range(len(specs)):
            e = plan.result(i).energies[0]
            e = e[~np.isnan(e)]
            expl.append(1.0 - e[-1] / e[0])
            removed.append(e[0] - e[-1])
        out[name] = {"seconds": dt, "atoms": atoms, "atoms_per_s": atoms / dt,
                     "entries": plan.entries_evaluated(), "explained": np.array(expl),
                     "removed": float(np.sum(removed))}
        if name == "proposal":
            out[name]["stats"] = plan.proposal_stats()
        plan.close()
        progress(f"proposal leg: {name} {atoms / dt:.0f} atoms/s")
    d = (out["plain"]["explained"] - out["proposal"]["explained"]) / out["plain"]["explained"]
    st = out["proposal"]["stats"]
    return {
        "workload": (f"uneven batch: {len(specs)} single-channel signals of 1024..16384 samples at 256 Hz, "
                     f"pink noise + Poisson(6) atoms, energyError 0.01, 99% / at most 100 iterations"),
        "network": "empi_amd/proposal_net.npy" if net is not None else None,
        "top_k": 8, "delta": args.proposal_delta,
        "plain_seconds": out["plain"]["seconds"], "proposal_seconds": out["proposal"]["seconds"],
        "atoms": out["proposal"]["atoms"],
        "plain_atoms_per_s": out["plain"]["atoms_per_s"],
        "proposal_atoms_per_s": out["proposal"]["atoms_per_s"],
        "speedup": out["proposal"]["atoms_per_s"] / out["plain"]["atoms_per_s"],
        "entries_ratio": out["plain"]["entries"] / out["proposal"]["entries"],
        "explained_deficit_max": float(d.max()), "explained_deficit_mean": float(d.mean()),
        "explained_deficit_batch": (out["plain"]["removed"] - out["proposal"]["removed"]) / out["plain"]["removed"],
        "fallback_regions_per_unit_iteration": st["fallback_regions"] / max(1, st["certified_iterations"]),
        "dirty_regions_per_unit_iteration": st["dirty_checked"] / max(1, st["certified_iterations"]),
    }


def cpu_sample(x, fs, eps, variant, iters, percent, limit, what):
    """FP64 oracle (OpenMP over scales, own FFT) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = min(16, os.cpu_count() or 1)
    d = po.Dictionary(fs, x.shape[-1], eps)
    t0 = time.perf_counter()
    r = po.decompose(d, np.asarray(x, np.float64), variant, iters, percent, threads=threads, iter_limit=limit)
    dt = time.perf_counter() - t0
    atoms = int(r.counts.sum())
    return {"value": atoms / dt, "unit": "atoms/s", "cores": threads, "kind": "port",
            "sample": f"{what}: {atoms} atoms in {dt:.1f} s, FP64 oracle with its own radix-2 FFT",
            "atoms": atoms, "seconds": dt}


def workload(args):
    """(specs, per-item signals, description, CPU-baseline sampler)"""
    import empi_amd as E
    from empi_amd import synth
    if args.workload == "smp-batch":
        fs, N, C, items = 256.0, 5120, args.channels, args.epochs
        I = args.iterations or 100
        sig = synth.batched_channels(items=items * C, N=N, fs=fs, atoms=8)
        specs = [E.ItemSpec(C, N, fs, E.SMP, I, 99.0, 0.01) for _ in range(items)]
        signals = [sig[i * C:(i + 1) * C] for i in range(items)]
        desc = (f"SMP batched plain method: {C} channels x {items} epochs x {N} float32 samples at 256 Hz, "
                f"pink noise + 8 Gabor atoms, energyError 0.01, {I} iterations, energyPercent 99")

        def cpu():
            return cpu_sample(signals[0][:1], fs, 0.01, 0, I, 99.0, args.cpu_iterations,
                              f"1 of the channel-epochs (5120 samples), full map build + first "
                              f"{args.cpu_iterations} iterations")
        return specs, signals, desc, cpu
    if args.workload == "single":
        I = args.iterations or 50
        x = synth.single_channel()
        specs = [E.ItemSpec(1, 1280, 128.0, E.SMP, I, 99.0, 0.01)]
        desc = (f"SMP single channel, 1280 float32 samples at 128 Hz, 5 atoms + noise 0.1, energyError 0.01, "
                f"{I} iterations, energyPercent 99")

        def cpu():
            return cpu_sample(x[None, :], 128.0, 0.01, 0, I, 99.0, I, "the whole decomposition")
        return specs, [x[None, :]], desc, cpu
    if args.workload == "mmp3":
        I = args.iterations or 200
        x = synth.shared_multichannel()
        specs = [E.ItemSpec(32, 10240, 512.0, E.MMP3, I, 99.0, 0.005)]
        desc = (f"MMP3: 32 channels x 10240 float32 samples at 512 Hz, 12 shared Gabor atoms + noise 0.2, "
                f"energyError 0.005, {I} iterations")

        def cpu():
            return cpu_sample(x, 512.0, 0.005, 3, I, 99.0, 1,
                              "full map build + first iteration (FP64 maps of 32 channels)")
        return specs, [x], desc, cpu
    I = args.iterations or 100
    xs = synth.uneven_batch(items=args.items)
    specs = [E.ItemSpec(1, len(x), 256.0, E.SMP, I, 99.0, 0.01) for x in xs]
    desc = (f"Uneven batch: {args.items} single-channel float32 signals of 1024..16384 samples at 256 Hz, "
            f"pink noise + Poisson(6) atoms, energyError 0.01, 99% / at most {I} iterations")

    def cpu():
        atoms, secs, cores = 0, 0.0, 1
        for x in xs[:4]:
            r = cpu_sample(x[None, :], 256.0, 0.01, 0, I, 99.0, I, "")
            atoms += r["atoms"]
            secs += r["seconds"]
            cores = r["cores"]
        return {"value": atoms / secs, "unit": "atoms/s", "cores": cores, "kind": "port",
                "sample": f"first 4 signals, whole decompositions ({atoms} atoms, {secs:.1f} s), "
                          f"FP64 oracle with its own radix-2 FFT"}
    return specs, [x[None, :] for x in xs], desc, cpu


def measured_traffic(kernel_q, desc):
    """HBM bytes per launch of the dominant column kernel from the committed
    FETCH_SIZE / WRITE_SIZE summary of the same workload, else None."""
    import glob
    for path in glob.glob(os.path.join(ROOT, "profiles", "k_columns_*_traffic_*.json")):
        with open(path) as f:
            d = json.load(f)
        if d.get("sub_length") == kernel_q and d.get("workload") == desc:
            return d["hbm_bytes_per_launch"], os.path.relpath(path, ROOT)
    return None, None


def measured_valu(kernel_q, desc):
    """VALU lane-instructions per map entry of the dominant column class from
    the committed SQ_INSTS_VALU summary of the same workload, else None."""
    path = os.path.join(ROOT, "profiles", "k_columns_valu_mix_smp_1280x5120.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    per = d["valu_lane_instr_per_entry"].get(str(kernel_q))
    if d.get("workload") != desc or per is None:
        return None
    return per, d["peak_wave_instr_per_s"], os.path.relpath(path, ROOT)


def copy_calibration():
    """One device-to-device copy of exactly 1 GiB: the reference reading for
    the FETCH_SIZE / WRITE_SIZE counters (expected 1048576 KiB each)."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4}))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "copy-calibration":
        copy_calibration()
        return
    import torch
    import empi_amd as E

    if args.workload == "proposal":
        prop = proposal_leg(args, torch)
        print(json.dumps({
            "metric": METRIC, "value": prop["proposal_atoms_per_s"], "unit": "atoms/s", "n_gpus": 1,
            "steps": 1, "warmup": 1, "ms_per_step": prop["proposal_seconds"] * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32 maps / fp64 energy", "data": "synthetic",
            "config": {"workload": prop["workload"], "method": "proposal"}, "proposal": prop}))
        return

    specs, signals, desc, cpu = workload(args)
    plan = E.Plan(specs)
    flat = np.concatenate([np.asarray(s, np.float32).ravel() for s in signals])
    progress(f"plan ready: {len(specs)} items, {flat.size} samples")
    dev = torch.from_numpy(flat).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()

    if args.tuning:
        plan.set_tuning([int(v) for v in args.tuning.split(",")])
        tuning = plan.tuning()
    else:
        t_tune = time.perf_counter()
        tuning = plan.autotune(device_ptr=ptr, repeats=2)
        progress(f"autotuned column kernels in {time.perf_counter() - t_tune:.1f} s")

    def step():
        plan.run(device_ptr=ptr)

    for i in range(args.warmup):
        step()
        progress(f"warmup step {i + 1} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        progress(f"timed step {i + 1} done")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    atoms = 0
    explained = []
    for i in range(len(specs)):
        r = plan.result(i)
        if r.status != 0:
            raise RuntimeError(f"item {i}: {E.STATUS_NAMES.get(r.status)}")
        atoms += int(r.counts.sum())
        for e in r.energies:
            last = e[~np.isnan(e)][-1]
            explained.append(1.0 - last / e[0])
    value = atoms * args.steps / dt

    # per-class kernel time with HIP events on the launch stream (one profiled step)
    plan.set_profiling(True)
    step()
    plan.set_profiling(False)
    progress("profiled step done")
    stats = plan.kernel_stats()
    dom = max(stats, key=lambda s: s["total_ms"])
    # algorithmic bytes: one FP32 complex map value per channel + one FP32 product per entry
    bytes_per_launch = (8 * dom["entry_channels"] + 4 * dom["entries"]) / dom["launches"]
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(dom["sub_length"], desc)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                "kernel": f"k_columns<{dom['sub_length']}>", "avg_launch_ms": avg_ms,
                "entries_per_launch": dom["entries"] / dom["launches"]}
    # the bound that binds: VALU issue (wave instructions per second against
    # 256 CUs x 4 SIMDs x one wave64 instruction per 4 cycles)
    valu = measured_valu(dom["sub_length"], desc)
    if valu:
        per_entry, peak, src = valu
        rate = roofline["entries_per_launch"] * per_entry / 64 / (avg_ms * 1e-3)
        roofline["valu"] = {"lane_instr_per_entry": per_entry, "wave_instr_per_s": rate,
                            "peak_wave_instr_per_s": peak, "issue_frac": rate / peak, "source": src}

    prop = None
    if args.workload == "smp-batch" and not args.no_proposal:
        prop = proposal_leg(args, torch)

    cpu_res = None
    if not args.no_cpu_baseline:
        progress("cpu baseline sample")
        cpu_res = cpu()
        cpu_res.pop("atoms", None)
        cpu_res.pop("seconds", None)

    out = {
        "metric": METRIC,
        "value": value, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 maps / fp64 energy", "data": "synthetic",
        "config": {"workload": desc, "method": "plain", "items": len(specs), "atoms_per_step": atoms,
                   "mean_explained_energy": float(np.mean(explained)),
                   "autotuned_variants": [int(v) for v in tuning]},
        "roofline": roofline,
        "cpu_baseline": cpu_res,
        "proposal": prop,
        "kernels": stats,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
