"""CPU baseline of the decomposition rate: the FP64 oracle (the reference's
loop restated, OpenMP over scales as gabor.cpp:77-82, 89-107, 198-208, its own
radix-2 FFT) timed on host cores.

TEST INFRASTRUCTURE ONLY: bench.py runs this file as a child process for its
cpu_baseline leg (so that OMP_NUM_THREADS and OMP_PROC_BIND reach a fresh
OpenMP runtime); nothing in the product path imports or runs it.

    python oracle/cpu_baseline.py uneven [--iterations 10] [--repeats 3]
    python oracle/cpu_baseline.py smp-batch ...
    python oracle/cpu_baseline.py check --cases cases.npz

The timing modes print one JSON object: the timed sample, median seconds and
atoms/s over the repeats with the inherited OMP_NUM_THREADS, a build-only run
(full map build + first atom) with the same threads and with one thread, and
the whole-decomposition rate extrapolated from the measured build and
per-iteration times.  `check` runs the oracle on engine results (the
checker of the benchmark's oracle-error block): for each case the oracle
follows the engine through near-ties and every iteration is compared by the
parity rules of tests/parity.py.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import pyoracle as po  # noqa: E402
from empi_amd import synth  # noqa: E402


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def timed(d, x, threads, limit, iters=100, percent=99.0):
    t0 = time.perf_counter()
    r = po.decompose(d, x, po.SMP, iters, percent, threads=threads, iter_limit=limit)
    return time.perf_counter() - t0, int(r.counts.sum())


def check(path):
    """Errors of engine results against the oracle following them."""
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))
    import parity
    z = np.load(path)
    n = 1 + max(int(k.split(":")[0]) for k in z.files)
    out = []
    for i in range(n):
        c = {k.split(":", 1)[1]: z[k] for k in z.files if k.startswith(f"{i}:")}
        x = np.asarray(c["signal"], np.float64)
        variant, iters, limit = int(c["variant"]), int(c["iters"]), int(c["limit"])
        d = po.Dictionary(float(c["fs"]), x.shape[1], float(c["eps"]))
        ora = parity.oracle_following(d, x, variant, iters, float(c["percent"]), c["index"], c["counts"],
                                      iter_limit=limit or None)
        rows = []
        for ch in range(x.shape[0]):
            r = ch if variant == po.SMP else 0
            err = parity.compare_rows(c["params"][ch], c["index"][ch], c["counts"][ch], ora.params[ch],
                                      ora.index[ch], ora.counts[ch], ora.followed[r], limited=limit > 0)
            err["energy_rel"] = parity.energy_rel(c["energies"][r], ora.energies[r], err["n"])
            rows.append(err)
        ok = all(e["index_mismatch"] == 0 and (e["limited"] or e["count_gpu"] == e["count_oracle"]) and
                 e["amp_rel"] <= parity.AMP_RTOL and e["mod_rel"] <= parity.AMP_RTOL and
                 e["phase_abs"] <= parity.PHASE_ATOL and e["energy_rel"] <= parity.ENERGY_RTOL for e in rows)
        out.append({"unit": str(c["name"]), "atoms_compared": sum(e["n"] for e in rows),
                    "iterations_compared": rows[0]["n"], "index_mismatches": sum(e["index_mismatch"] for e in rows),
                    "near_ties_followed": int(ora.followed.sum()),
                    "amplitude_rel": max(e["amp_rel"] for e in rows), "modulus_rel": max(e["mod_rel"] for e in rows),
                    "phase_abs": max(e["phase_abs"] for e in rows), "energy_rel": max(e["energy_rel"] for e in rows),
                    "within_contract": ok})
    summary = {"units": len(out), "atoms_compared": sum(o["atoms_compared"] for o in out),
               "index_mismatches": sum(o["index_mismatches"] for o in out),
               "amplitude_rel": max(o["amplitude_rel"] for o in out), "phase_abs": max(o["phase_abs"] for o in out),
               "energy_rel": max(o["energy_rel"] for o in out), "all_within_contract": all(o["within_contract"] for o in out),
               "contract": "indices exact where the oracle's top-two gap > 1e-4 (near-ties followed), amplitude and "
                           "modulus 1e-4 relative, phase 1e-4 pi, residual energy 1e-5 relative"}
    print(json.dumps({"cases": out, "summary": summary}, default=float))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["uneven", "smp-batch", "check"])
    ap.add_argument("--cases", default=None)
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--atoms-per-signal", type=float, default=100.0,
                    help="mean atoms per decomposition on the GPU, for the extrapolated rate")
    a = ap.parse_args()
    if a.workload == "check":
        check(a.cases)
        return
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if a.workload == "uneven":
        xs = synth.uneven_batch()
        i = int(np.argsort([len(s) for s in xs])[len(xs) // 2])
        x = xs[i][None, :].astype(np.float64)
        fs, eps = 256.0, 0.01
        what = f"uneven-batch signal {i} (the median length, {x.shape[1]} samples at 256 Hz)"
    else:
        x = synth.batched_channels(items=1, N=5120, fs=256.0)[:1].astype(np.float64)
        fs, eps = 256.0, 0.01
        what = "channel-epoch 0 of the 64 x 20 x 5120 batch"
    d = po.Dictionary(fs, x.shape[1], eps)
    runs = [timed(d, x, threads, a.iterations) for _ in range(a.repeats)]
    secs = float(np.median([t for t, _ in runs]))
    atoms = runs[0][1]
    build, _ = timed(d, x, threads, 1)
    build1, _ = timed(d, x, 1, 1)
    per_iter = max(0.0, (secs - build) / max(1, atoms - 1))
    n = a.atoms_per_signal
    print(json.dumps({
        "value": atoms / secs, "unit": "atoms/s", "cores": threads, "kind": "port",
        "sample": (f"{what}: full map build + first {atoms} atoms, median of {a.repeats} runs "
                   f"({secs:.2f} s), FP64 oracle with its own radix-2 FFT, OMP_PROC_BIND="
                   f"{os.environ.get('OMP_PROC_BIND', 'unset')}"),
        "seconds_median": secs, "seconds_runs": [t for t, _ in runs], "atoms": atoms,
        "build_seconds": build, "build_seconds_1_thread": build1,
        "thread_speedup_on_build": build1 / build if build > 0 else None,
        "per_iteration_seconds": per_iter,
        "extrapolated_whole_decomposition_atoms_per_s": n / (build + (n - 1) * per_iter),
        "extrapolation": (f"{n:.1f} atoms per signal (the GPU run's mean) after one full build: "
                          f"build {build:.2f} s + {n - 1:.1f} x {per_iter:.3f} s per iteration"),
        "nproc": os.cpu_count(), "cpu_model": cpu_model(), "fft": "own radix-2 (FFTW3 is absent)",
    }))


if __name__ == "__main__":
    main()
