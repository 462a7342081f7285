"""Benchmark: Matching Pursuit decomposition on one MI355X.

Headline (default, --workload proposal): the uneven batch with the learned
proposal against the plain method.  512 single-channel float32 signals of
1024..16384 samples at 256 Hz, each pink noise + Poisson(6) random Gabor
atoms; optimal Gabor dictionary with energyError 0.01, energyPercent 99 / at
most 100 iterations; FP32 maps, FP64 energy bookkeeping.  A step decomposes
the whole batch (full map build + MP loop) with the input resident in HBM,
the proposal network scoring 8 regions per iteration and the certified
fallback at delta 1e-4.  `value` is the proposal's atoms/s over the K timed
steps; the plain method's atoms/s (median of its runs), the speed-up and
the explained-energy deficit (batch and worst signal) sit beside it.

Secondary leg (--no-secondary skips it): the SMP batch, 64 channels x 20
epochs x 5120 samples at 256 Hz, pink noise + 8 atoms, 100 iterations, plain
method, median of 3 steps.

Both legs carry their dominant column kernel's roofline (HIP events on the
launch stream, one profiled extra step) and their error against the FP64
oracle on sampled units: the oracle follows the engine through near-ties
(within 1e-4 of its own best) and every iteration is compared.  The oracle
checks and the CPU baseline run in child processes (oracle/cpu_baseline.py),
the CPU baseline with OMP_PROC_BIND=close and the inherited OMP_NUM_THREADS.

Other workloads (one plan, plain method, K timed steps): smp-batch, single
(1280 samples at 128 Hz, 50 iterations), mmp3 (32 channels x 10240 samples
at 512 Hz, energyError 0.005, 200 iterations), uneven (the batch above,
plain).

Prints one JSON line on stdout (progress goes to stderr).  --gpus above 1 is
refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "atoms/sec to 99% energy on 1 MI355X; proposal vs plain at stated energy error"
UNEVEN_ID, SMP_ID = "uneven512_proposal", "smp_1280x5120"


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="proposal",
                    choices=["proposal", "smp-batch", "single", "mmp3", "uneven", "copy-calibration"])
    ap.add_argument("--items", type=int, default=512, help="signals of the uneven batch")
    ap.add_argument("--delta", type=float, default=1e-4)
    ap.add_argument("--top-k", type=int, default=8)
    ap.add_argument("--plain-runs", type=int, default=3, help="timed runs of the plain method (median)")
    ap.add_argument("--unbatched-sample", type=int, default=32,
                    help="signals of the one-plan-per-signal plain leg (0 skips it)")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--channels", type=int, default=64)
    ap.add_argument("--iterations", type=int, default=None)
    ap.add_argument("--no-secondary", action="store_true", help="skip the SMP-batch leg")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU baseline and the oracle checks")
    ap.add_argument("--cpu-iterations", type=int, default=10, help="atoms of the CPU-baseline sample")
    ap.add_argument("--tuning", default=None,
                    help="comma-separated column-kernel tuning table instead of autotuning")
    return ap.parse_args()


# ------------------------------------------------------------------ helpers

def run_steps(plan, ptr, torch, warmup, steps, what):
    """W untimed steps, then K steps bracketed by synchronisations; returns
    (total seconds, per-step seconds)."""
    for i in range(warmup):
        plan.run(device_ptr=ptr)
        progress(f"{what}: warmup step {i + 1} done")
    torch.cuda.synchronize()
    per = []
    t0 = time.perf_counter()
    for i in range(steps):
        t1 = time.perf_counter()
        plan.run(device_ptr=ptr)
        per.append(time.perf_counter() - t1)
        progress(f"{what}: timed step {i + 1} done ({per[-1]:.2f} s)")
    torch.cuda.synchronize()
    return time.perf_counter() - t0, per


def item_stats(plan, n):
    """(atoms, explained energy per row, energy removed per row) of the last run."""
    import empi_amd as E
    atoms, expl, removed = 0, [], []
    for i in range(n):
        r = plan.result(i)
        if r.status != 0:
            raise RuntimeError(f"item {i}: {E.STATUS_NAMES.get(r.status)}")
        atoms += int(r.counts.sum())
        for e in r.energies:
            e = e[~np.isnan(e)]
            expl.append(1.0 - e[-1] / e[0])
            removed.append(e[0] - e[-1])
    return atoms, np.array(expl), np.array(removed)


def profile_json(kind, sub_length, workload_id):
    """A committed profiles/*.json summary of this kernel class and workload."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"k_columns*_{kind}_*.json"))):
        with open(path) as f:
            d = json.load(f)
        if d.get("workload_id") == workload_id and (sub_length is None or d.get("sub_length") == sub_length):
            return d, os.path.relpath(path, ROOT)
    return None, None


def roofline(plan, ptr, workload_id):
    """Dominant column class of one profiled step (HIP events on the plan's
    stream): algorithmic bytes (8 C + 4 per re-evaluated entry) per launch over
    the average launch time, against the HBM peak; measured HBM bytes and the
    VALU issue rate from the committed rocprofv3 summaries of the workload."""
    plan.set_profiling(True)
    plan.run(device_ptr=ptr)
    plan.set_profiling(False)
    stats = plan.kernel_stats()
    dom = max(stats, key=lambda s: s["total_ms"])
    bytes_per_launch = (8 * dom["entry_channels"] + 4 * dom["entries"]) / dom["launches"]
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, src = profile_json("traffic", dom["sub_length"], workload_id)
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
           "traffic_source": src, "kernel": f"k_columns<{dom['sub_length']}>", "avg_launch_ms": avg_ms,
           "entries_per_launch": dom["entries"] / dom["launches"],
           "algorithmic_bytes_per_launch": bytes_per_launch,
           "share_of_kernel_time": dom["total_ms"] / sum(s["total_ms"] for s in stats)}
    valu, vsrc = profile_json("valu_mix", None, workload_id)
    if valu and str(dom["sub_length"]) in valu["valu_lane_instr_per_entry"]:
        per = valu["valu_lane_instr_per_entry"][str(dom["sub_length"])]
        rate = out["entries_per_launch"] * per / 64 / (avg_ms * 1e-3)
        out["valu"] = {"lane_instr_per_entry": per, "wave_instr_per_s": rate,
                       "peak_wave_instr_per_s": valu["peak_wave_instr_per_s"],
                       "issue_frac": rate / valu["peak_wave_instr_per_s"], "source": vsrc}
    return out, stats


def child(args_list, env=None, timeout=900):
    """Runs oracle/cpu_baseline.py in a child process; returns its JSON."""
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py")] + args_list
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(args_list[:1])} failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def oracle_errors(cases):
    """FP64-oracle errors of sampled units: cases = list of dicts (name, signal
    [C][N], fs, eps, variant, iters, percent, limit, index, counts, params,
    energies).  Written to a temporary file and checked in a child process."""
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cases.npz")
        flat = {}
        for i, c in enumerate(cases):
            for k, v in c.items():
                flat[f"{i}:{k}"] = np.asarray(v)
        np.savez(path, **flat)
        return child(["check", "--cases", path])


# ------------------------------------------------------------------ legs

def uneven_signals(args):
    from empi_amd import synth
    xs = synth.uneven_batch(items=args.items)
    return xs, (f"uneven batch: {len(xs)} single-channel float32 signals of 1024..16384 samples at 256 Hz, "
                f"pink noise + Poisson(6) Gabor atoms, energyError 0.01, energyPercent 99 / at most 100 "
                f"iterations")


def unbatched_plain(xs, ptr, torch, tuning, sample):
    """The plain method one signal at a time: one single-signal plan per
    signal of a length-stratified sample (`sample` signals evenly spaced in
    length order), run one after another with the same tuning table.  Plans
    are created and warmed up before the timed loop."""
    import empi_amd as E
    offs = np.concatenate([[0], np.cumsum([len(x) for x in xs])])
    order = np.argsort([len(x) for x in xs], kind="stable")
    pick = [int(order[j]) for j in np.linspace(0, len(xs) - 1, sample).round().astype(int)]
    plans = []
    for i in pick:
        p = E.Plan([E.ItemSpec(1, len(xs[i]), 256.0, E.SMP, 100, 99.0, 0.01)])
        p.set_tuning(tuning)
        plans.append((i, p, ptr + 4 * int(offs[i])))
    for _, p, a in plans:
        p.run(device_ptr=a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _, p, a in plans:
        p.run(device_ptr=a)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    atoms = sum(int(p.result(0).counts.sum()) for _, p, _ in plans)
    for _, p, _ in plans:
        p.close()
    progress(f"unbatched plain: {atoms} atoms of {len(pick)} signals in {sec:.2f} s")
    return {"atoms_per_s": atoms / sec, "atoms": atoms, "seconds": sec, "signals": len(pick),
            "sample": f"{len(pick)} signals evenly spaced in length order, one plan each, run in sequence",
            "samples_total": int(sum(len(xs[i]) for i in pick))}


def headline(args, torch):
    import empi_amd as E
    from empi_amd import proposal
    xs, desc = uneven_signals(args)
    specs = [E.ItemSpec(1, len(x), 256.0, E.SMP, 100, 99.0, 0.01) for x in xs]
    dev = torch.from_numpy(np.concatenate([x.astype(np.float32) for x in xs])).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()
    net = proposal.load()
    if net is None:
        raise RuntimeError("empi_amd/proposal_net.npy missing")

    plain = E.Plan(specs)
    if args.tuning:
        plain.set_tuning([int(v) for v in args.tuning.split(",")])
        tuning = plain.tuning()
    else:
        t0 = time.perf_counter()
        tuning = plain.autotune(device_ptr=ptr, repeats=2)
        progress(f"autotuned column kernels in {time.perf_counter() - t0:.1f} s")
    _, plain_per = run_steps(plain, ptr, torch, 1, args.plain_runs, "plain")
    plain_atoms, plain_expl, plain_removed = item_stats(plain, len(specs))
    plain_res = [plain.result(i) for i in range(len(specs))]
    plain_entries = plain.entries_evaluated()
    plain.close()
    unbatched = unbatched_plain(xs, ptr, torch, tuning, args.unbatched_sample) if args.unbatched_sample else None

    prop = E.Plan(specs)
    prop.set_proposal(E.PROPOSAL_ON, net, top_k=args.top_k, delta=args.delta)
    if args.tuning:
        prop.set_tuning(tuning)
        prop_tuning = tuning
    else:
        # the proposal runs the low-bin kernel variants: tuned on its own plan
        t0 = time.perf_counter()
        prop_tuning = prop.autotune(device_ptr=ptr, repeats=2)
        progress(f"autotuned the proposal's column kernels in {time.perf_counter() - t0:.1f} s")
    total, per = run_steps(prop, ptr, torch, args.warmup, args.steps, "proposal")
    atoms, expl, removed = item_stats(prop, len(specs))
    value = atoms * args.steps / total
    stats = prop.proposal_stats()
    entries = prop.entries_evaluated()
    res = [prop.result(i) for i in range(len(specs))]
    roof, kernels = roofline(prop, ptr, UNEVEN_ID)
    prop.close()
    plain_rate = plain_atoms / float(np.median(plain_per))
    deficit = (plain_expl - expl) / plain_expl
    line = {
        "value": value, "ms_per_step": total / args.steps * 1e3,
        "step_seconds": per,
        "config": {"workload": desc, "workload_id": UNEVEN_ID, "method": "proposal",
                   "network": "empi_amd/proposal_net.npy", "top_k": args.top_k, "delta": args.delta,
                   "items": len(specs), "atoms_per_step": atoms,
                   "autotuned_variants": [int(v) for v in prop_tuning],
                   "plain_autotuned_variants": [int(v) for v in tuning]},
        "plain": {"atoms_per_s": plain_rate, "atoms_per_run": plain_atoms, "run_seconds": plain_per,
                  "median_of": len(plain_per), "entries": plain_entries},
        "proposal_vs_plain": {
            "speedup": value / plain_rate,
            "entries_ratio": plain_entries / entries,
            "explained_deficit_batch": float((plain_removed.sum() - removed.sum()) / plain_removed.sum()),
            "explained_deficit_max": float(deficit.max()),
            "explained_deficit_mean": float(deficit.mean()),
            "fallback_regions_per_unit_iteration": stats["fallback_regions"] / max(1, stats["certified_iterations"]),
            "dirty_regions_per_unit_iteration": stats["dirty_checked"] / max(1, stats["certified_iterations"]),
        },
        "roofline": roof, "kernels": kernels,
    }
    if unbatched:
        line["unbatched_plain"] = {**unbatched,
                                   "batched_proposal_vs_unbatched": value / unbatched["atoms_per_s"],
                                   "batched_plain_vs_unbatched": plain_rate / unbatched["atoms_per_s"]}
    # oracle samples: the two shortest signals whole, the median-length one over 20 iterations
    order = np.argsort([len(x) for x in xs])
    cases = []
    for i, limit in [(int(order[0]), 0), (int(order[1]), 0), (int(order[len(xs) // 2]), 20)]:
        r = res[i]
        cases.append({"name": f"signal {i} ({len(xs[i])} samples)", "signal": xs[i][None, :], "fs": 256.0,
                      "eps": 0.01, "variant": 0, "iters": 100, "percent": 99.0, "limit": limit,
                      "index": r.index, "counts": r.counts, "params": r.params, "energies": r.energies})
    return line, cases, float(np.mean([int(r.counts.sum()) for r in res]))


def smp_batch(args, torch, steps, warmup=1, tuning=None):
    import empi_amd as E
    from empi_amd import synth
    fs, N, C, items = 256.0, 5120, args.channels, args.epochs
    I = args.iterations or 100
    sig = synth.batched_channels(items=items * C, N=N, fs=fs, atoms=8)
    specs = [E.ItemSpec(C, N, fs, E.SMP, I, 99.0, 0.01) for _ in range(items)]
    desc = (f"SMP batched plain method: {C} channels x {items} epochs x {N} float32 samples at 256 Hz, "
            f"pink noise + 8 Gabor atoms, energyError 0.01, {I} iterations, energyPercent 99")
    plan = E.Plan(specs)
    dev = torch.from_numpy(np.ascontiguousarray(sig, np.float32).ravel()).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()
    if tuning:
        plan.set_tuning([int(v) for v in tuning.split(",")])
        tuning = plan.tuning()
    else:
        tuning = plan.autotune(device_ptr=ptr, repeats=2)
    total, per = run_steps(plan, ptr, torch, warmup, steps, "smp-batch")
    atoms, expl, _ = item_stats(plan, len(specs))
    wid = SMP_ID if (C, items, I) == (64, 20, 100) else None
    roof, kernels = roofline(plan, ptr, wid)
    r = plan.result(0)
    case = {"name": "epoch 0 channel 0", "signal": sig[0][None, :], "fs": fs, "eps": 0.01, "variant": 0,
            "iters": I, "percent": 99.0, "limit": 20, "index": r.index[:1], "counts": r.counts[:1],
            "params": r.params[:1], "energies": r.energies[:1]}
    plan.close()
    return {"workload": desc, "workload_id": wid, "atoms_per_s": atoms / float(np.median(per)),
            "atoms_per_step": atoms, "step_seconds": per, "median_of": len(per),
            "mean_explained_energy": float(np.mean(expl)), "autotuned_variants": [int(v) for v in tuning],
            "roofline": roof, "kernels": kernels}, case


def single_plan(args, torch):
    """--workload single | mmp3 | uneven: one plan, plain method."""
    assert args.workload in ("single", "mmp3", "uneven"), args.workload
    import empi_amd as E
    from empi_amd import synth
    if args.workload == "single":
        I = args.iterations or 50
        specs = [E.ItemSpec(1, 1280, 128.0, E.SMP, I, 99.0, 0.01)]
        sig = [synth.single_channel()[None, :]]
        desc = (f"SMP single channel, 1280 float32 samples at 128 Hz, 5 atoms + noise 0.1, energyError 0.01, "
                f"{I} iterations, energyPercent 99")
    elif args.workload == "mmp3":
        I = args.iterations or 200
        specs = [E.ItemSpec(32, 10240, 512.0, E.MMP3, I, 99.0, 0.005)]
        sig = [synth.shared_multichannel()]
        desc = (f"MMP3: 32 channels x 10240 float32 samples at 512 Hz, 12 shared Gabor atoms + noise 0.2, "
                f"energyError 0.005, {I} iterations")
    else:
        xs, desc = uneven_signals(args)
        specs = [E.ItemSpec(1, len(x), 256.0, E.SMP, 100, 99.0, 0.01) for x in xs]
        sig = [x[None, :] for x in xs]
        desc += ", plain method"
    plan = E.Plan(specs)
    dev = torch.from_numpy(np.concatenate([np.asarray(s, np.float32).ravel() for s in sig])).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()
    if args.tuning:
        plan.set_tuning([int(v) for v in args.tuning.split(",")])
        tuning = plan.tuning()
    else:
        tuning = plan.autotune(device_ptr=ptr, repeats=2)
    total, per = run_steps(plan, ptr, torch, args.warmup, args.steps, args.workload)
    atoms, expl, _ = item_stats(plan, len(specs))
    roof, kernels = roofline(plan, ptr, None)
    return {"value": atoms * args.steps / total, "ms_per_step": total / args.steps * 1e3, "step_seconds": per,
            "config": {"workload": desc, "method": "plain", "items": len(specs), "atoms_per_step": atoms,
                       "mean_explained_energy": float(np.mean(expl)),
                       "autotuned_variants": [int(v) for v in tuning]},
            "roofline": roof, "kernels": kernels}


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

    base = {"metric": METRIC, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32 maps / fp64 energy", "data": "synthetic"}
    if args.workload == "smp-batch":
        sec, _ = smp_batch(args, torch, args.steps, args.warmup, args.tuning)
        line = {"value": sec["atoms_per_s"], "ms_per_step": float(np.median(sec["step_seconds"])) * 1e3,
                "step_seconds": sec["step_seconds"],
                "config": {"workload": sec["workload"], "workload_id": sec["workload_id"], "method": "plain",
                           "items": args.epochs, "atoms_per_step": sec["atoms_per_step"],
                           "mean_explained_energy": sec["mean_explained_energy"],
                           "autotuned_variants": sec["autotuned_variants"]},
                "roofline": sec["roofline"], "kernels": sec["kernels"]}
        print(json.dumps({**base, **line}))
        return
    if args.workload != "proposal":
        line = single_plan(args, torch)
        print(json.dumps({**base, **line}))
        return

    line, cases, atoms_per_signal = headline(args, torch)
    secondary = None
    if not args.no_secondary:
        secondary, case = smp_batch(args, torch, 3)
        progress(f"secondary leg: {secondary['atoms_per_s']:.0f} atoms/s")
    oracle = cpu = None
    if not args.no_cpu_baseline:
        progress("oracle checks of sampled units")
        checks = oracle_errors(cases + ([case] if secondary else []))
        oracle = {"headline": checks["cases"][:len(cases)], "summary": checks["summary"]}
        if secondary:
            secondary["oracle_error"] = checks["cases"][len(cases):]
        progress("cpu baseline")
        env = dict(os.environ, OMP_PROC_BIND="close")
        env.setdefault("OMP_NUM_THREADS", str(os.cpu_count() or 1))
        cpu = child(["uneven", "--iterations", str(args.cpu_iterations),
                     "--atoms-per-signal", f"{atoms_per_signal:.2f}"], env=env)
    print(json.dumps({**base, **line, "oracle_error": oracle, "cpu_baseline": cpu, "secondary": secondary}))


if __name__ == "__main__":
    main()
