"""Benchmark: GPU CLE throughput in cell*gene*steps/s on one MI355X.

Workload (one "step" of this benchmark = one full simulate() pass): a seeded
G=20 mixed-regulation network with telegraph promoters on every gene,
100,000 cells, 5000 Euler-Maruyama steps at dt=0.01, a snapshot every 500
steps, FP64 moment sums per snapshot and LogNormal-Binomial capture on the
final state.  Inputs (the compiled network) are resident in HBM before the
timed region; outputs stay in HBM.

Prints one JSON line.  Adds:
  roofline      VALU issue roofline of the stepper kernel (the dominant kernel),
                timed with HIP events on its launch stream;
  cpu_baseline  the FP64 CPU restatement (oracle/) timed on this host's cores
                on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "syntheticscrnaseq.jl_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

# MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz = 39.3e12 VALU lane-ops/s
# (the published 157.3 TFLOP/s FP32 peak is this x 2 (FMA) x 2 (packed)).
VALU_PEAK_TOPS = 39.3216
FP32_PEAK_TFLOPS = 157.3


def valu_census(net) -> float:
    """Algorithmic VALU lane-ops per cell*gene*step for a network (fixed rule).

    E = 36 FP32 ops for the update (cle.jl:79-120, every +,-,*,/,max,sqrt,pow = 1);
    R = 2 * additive in-edges per gene + 3 * gate source incidences per gene;
    telegraph = 3 ops; quarter-rate surcharge +3 for each of 2 sqrt and 1 rcp;
    Box-Muller for one normal pair (log, sqrt, sin, cos at 4 each + 5) = 21;
    Philox4x32-10 = 138 ops per call (20 32x32 products at 4 + 58 xor/add),
    0.5 call per gene for the two normals, +0.25 with telegraph switching.
    """
    G = net.G
    A = np.array(net.interactions)
    np.fill_diagonal(A, 0.0)
    d_add = np.count_nonzero(A) / G
    inc = sum(len(e.sources) for e in list(net.cooperative) + list(net.redundant)) / G
    tele = bool(np.any(np.isfinite(net.k_on)))
    ops = 36 + 2 * d_add + 3 * inc + 9 + 21 + 138 * 0.5
    if tele:
        ops += 3 + 138 * 0.25
    return float(ops)


def cpu_baseline(wl, target_s: float = 12.0) -> dict:
    """Time the FP64 oracle (a C restatement of cle.jl) on a bounded cell sample."""
    from oracle import oracle as orc

    threads = max(1, min(16, os.cpu_count() or 1))
    net, kin = wl["net"], wl["kin"]
    steps = int(np.ceil(wl["T"] / wl["dt"]))
    probe_cells = 4 * threads
    t0 = time.perf_counter()
    orc.cle(net, kin, dt=wl["dt"], n_steps=steps, cell_num=wl["cells"], cell_lo=0, cell_hi=probe_cells,
            seed=1, threads=threads)
    dt_probe = time.perf_counter() - t0
    rate = probe_cells / max(dt_probe, 1e-6)
    cells = int(min(wl["cells"], max(probe_cells, rate * target_s)))
    cells = max(threads, (cells // threads) * threads)
    t0 = time.perf_counter()
    orc.cle(net, kin, dt=wl["dt"], n_steps=steps, cell_num=wl["cells"], cell_lo=0, cell_hi=cells, seed=1,
            threads=threads)
    el = time.perf_counter() - t0
    cgs = float(cells) * net.G * steps
    return {"value": cgs / el, "unit": "cell*gene*steps/s", "cores": threads, "kind": "port",
            "sample": f"{cells} of {wl['cells']} cells x G={net.G} x {steps} steps (full horizon), "
                      f"FP64 C restatement of cle.jl with telegraph + AND/OR, OpenMP over cells, "
                      f"{el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=None, help="override cell count (testing only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(cells and networks would shard across GPUs with no exchange; see DESIGN.md).",
              file=sys.stderr)
        return 2

    import scrna_amd as sc
    from scrna_amd import samplers

    wl = samplers.gpu_g20_workload()
    if a.cells:
        wl["cells"] = a.cells
    net, kin = wl["net"], wl["kin"]
    plan = sc.Plan([net], sc.CLE(wl["dt"]), kin, cell_num=wl["cells"], T=wl["T"], readout="mrna",
                   capture=wl["capture"], seed=20251203, init=wl["init"], snapshot_every=wl["snapshot_every"])
    steps_per_run = sc.n_steps(wl["T"], wl["dt"])
    cgs_per_run = float(net.G) * wl["cells"] * steps_per_run

    for _ in range(a.warmup):
        plan.launch()
    plan.sync()
    print(f"# warmup done ({a.warmup}), timing {a.steps} runs", file=sys.stderr, flush=True)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        plan.launch()
        ms, _ = plan.stepper_ms()          # waits for this run's stepper kernel
        kernel_ms.append(ms)
    plan.sync()
    el = time.perf_counter() - t0

    res = plan.download(0)
    if not np.all(np.isfinite(res.y)):
        print("non-finite state", file=sys.stderr)
        return 1

    value = cgs_per_run * a.steps / el
    k_ms = float(np.mean(kernel_ms))
    census = valu_census(net)
    achieved = census * cgs_per_run / (k_ms * 1e-3) / 1e12
    out = {
        "metric": "cell*gene*steps/s (GPU CLE, FP32 state)",
        "value": value,
        "unit": "cell*gene*steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded G=20 network, Philox noise)",
        "config": {"workload": "G=20 mixed AND/OR network, telegraph on all genes, 100000 cells, "
                               "5000 steps dt=0.01, snapshot every 500, capture on last",
                   "G": net.G, "cells": wl["cells"], "sim_steps": steps_per_run,
                   "snapshot_every": wl["snapshot_every"], "capture": "LogNormal(0.1, 0.3)-Binomial",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s",
                     "frac": achieved / VALU_PEAK_TOPS, "traffic": None,
                     "kernel": "cle_stepper_kernel", "kernel_ms": k_ms, "census_ops_per_cgs": census,
                     "fp32_equiv": {"achieved": 4 * achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s"}},
    }
    if not a.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(wl)
    plan.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
