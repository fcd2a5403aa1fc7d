"""Benchmark: GPU CLE throughput in cell*gene*steps/s on one MI355X.

Default (--workload headline): the metric "BF16-MFMA vs FP32 CLE at <=1%
moment error; p99 latency", measured in one run.
  * value: the G=256 precision workload (additive network with 25% dense A,
    8 AND + 8 OR edges, 65,536 cells, 2000 Euler-Maruyama steps at dt=0.005)
    with the regulation contraction in BF16 on MFMA and FP32 state; one timed
    "step" = one full simulate() pass, inputs resident in HBM.
  * precision: the same workload with BF16x3 split precision, the fused
    FP32 forms and the reference-structured FP32 form (one kernel per
    broadcast, two SGEMMs per step), their speedups, and per-gene mean and
    variance errors against an FP64 CPU replay of a cell subset drawing the
    identical Philox stream.
  * g20: the G=20 telegraph workload (100,000 cells, 5000 steps, snapshots
    every 500, capture), the FP32 stepper's throughput and roofline.
  * serving: Poisson arrivals at a fixed QPS, p50 / p99 latency against the
    stated bound, aggregate throughput.
  * roofline: the dominant kernel of `value` (the dense MFMA stepper), timed
    with HIP events on its launch stream.
  * cpu_baseline / cpu_baselines: CPU restatements of the reference CLE
    (oracle/cpu_ref.py: cle.jl's cell-vectorised form with OpenBLAS products)
    and CLEFast's per-cell form (the FP64 oracle, OpenMP over cells), at one
    thread and at all cores used, on the G=5 workload, a G=20 slice and a G=256
    slice, with the host CPU named.

Other workloads (--workload): g20, g256, batch (4096 uneven networks with
population dynamics in one hipGraph), serve, tauleap (binomial tau-leap on the
G=20 network), sbi (networks drawn from the prior, simulated in one plan),
copy (plain device copy of known size, for scaling hardware byte counters).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "syntheticscrnaseq.jl_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

# MI355X published peaks: FP32 vector 157.3 TFLOP/s (256 CUs x 4 SIMD x 32 FP32
# lanes per clock x 2 (FMA) x 2.4 GHz), dense BF16 MFMA 2.5 PFLOP/s.
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0

# gfx950 VALU issue cost of each instruction class relative to one FP32 FMA,
# measured once on an MI355X with dependent chains of each instruction (8
# independent chains per lane, 8 waves per SIMD; table in
# profiles/valu_rates_gfx950.md).  Fixed: the census does not change between runs.
COST_FP32 = 1.0       # v_add/mul/fma/max_f32
COST_INT = 1.66       # 32-bit integer and logic (xor, add, shifts, compares)
COST_MUL32 = 1.54     # one 32x32-bit product (v_mad_u64_u32, v_mul_hi_u32)
COST_TRANS = 3.63     # v_log/exp/sqrt/rcp/sin/cos_f32
# FP64 and other classes, from the same microbenchmark (the tau-leap's executed mix):
COST_F64_FMA = 1.74   # v_fma_f64
COST_F64_ADD = 1.60   # v_add_f64
COST_F64_MUL = 1.65   # v_mul_f64
COST_F64_TRANS = 6.08  # v_rcp/sqrt/rsq_f64
COST_CVT = 1.53       # one f64 <-> i32 conversion
COST_OTHER = 1.56     # any other 32-bit VALU instruction (moves, selects, logic, compares; v_perm_b32 measured)


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


def valu_census(net, dense: bool = False) -> dict:
    """Algorithmic VALU work per cell*gene*step of a network (fixed operation
    counts), priced in FP32-FMA issue units with the measured gfx950 costs.

    Operation counts (the fewest per-lane instructions the algorithm needs on
    gfx950, so that the fraction of peak is a speed-of-light fraction):
      update E = 36 ops (cle.jl:79-120, every +,-,*,/,max,sqrt,pow = 1), of
        which 2 sqrt and 1 reciprocal are transcendental and 33 FP32; the
        steppers need 5 fewer (28 FP32): counts are never negative at the
        start of a step, so max(m, 0), max(p, 0) and the protein variance's
        max(., 0) are no-ops, the decay propensity mu_p*p serves drift and
        variance, and the basal rate is the regulation sum's first addend
        (the dense kernel's bias);
      regulation R = 2 * additive in-edges per gene (0 on the MFMA path, whose
        contraction is MFMA work) + 3 * gate source incidences per gene, FP32;
      Box-Muller for the pair of normals: log, sin, cos + 5 FP32 (its radius
        square root folds into the update's two, sqrt(r^2 * var));
      Philox4x32-10: per round two 32x32-bit products and two 3-input xors
        (v_bitop3_b32); the key schedule is wave-uniform, scalar work.  So 20
        products + 20 integer ops per call, 0.5 call per gene;
      telegraph switching: 6 integer ops per gene (its 18 switching bits are
        the spare low bits of its two normal words: 3 ops to join them, 1
        threshold select, 2 compares); the extra draw for the 2^-18 of cells
        whose 18 bits do not decide is not counted.
    Returns the class counts, the FMA-equivalent total and the FP32-equivalent
    flops (2 per FMA-equivalent op).
    """
    G = net.G
    A = np.array(net.interactions)
    np.fill_diagonal(A, 0.0)
    d_add = 0.0 if dense else np.count_nonzero(A) / G
    inc = sum(len(e.sources) for e in list(net.cooperative) + list(net.redundant)) / G
    tele = bool(np.any(np.isfinite(net.k_on)))
    calls = 0.5
    counts = {"fp32": 28 + 2 * d_add + 3 * inc + 5, "trans": 3 + 3,
              "int": 20 * calls + (6 if tele else 0), "mul32": 20 * calls}
    fma_eq = (counts["fp32"] * COST_FP32 + counts["trans"] * COST_TRANS + counts["int"] * COST_INT
              + counts["mul32"] * COST_MUL32)
    return {"counts": {k: round(v, 3) for k, v in counts.items()}, "fma_equiv_ops": fma_eq,
            "fp32_equiv_flops": 2.0 * fma_eq}


def host_cpu() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count() or 1}


def _threads_all() -> int:
    # a one-GPU host share is 16 cores, whatever os.cpu_count() reports
    return max(1, min(16, os.cpu_count() or 1))


def _sized(run, n_max: int, probe: int, target_s: float, quantum: int = 1):
    """Time run(probe), then run(n) with n (cells or steps) sized to about target_s seconds."""
    t0 = time.perf_counter()
    run(probe)
    rate = probe / max(time.perf_counter() - t0, 1e-6)
    n = int(min(n_max, max(probe, rate * target_s)))
    n = max(quantum, (n // quantum) * quantum)
    t0 = time.perf_counter()
    run(n)
    return n, time.perf_counter() - t0


def time_restatement(name: str, case, threads: int, target_s: float) -> dict:
    """Time one CPU restatement on a cell sample of one workload case
    (label, net, kinetics, dt, steps, cells of the workload, init)."""
    from oracle import cpu_ref
    from oracle import oracle as orc

    label, net, k, dt, steps, cells_max, init = case
    if name == "cle_ref":
        # cell-vectorised: up to 16384 of the workload's cells at once (its per-step cost is linear
        # in cells from there on), a sample of its steps
        cells = min(cells_max, 16384)

        def run(n):
            cpu_ref.cle_ref(net, k, dt=dt, n_steps=n, cells=cells, seed=1, init=init, threads=threads, noise="randn")
        steps, el = _sized(run, steps, 4, target_s)
    else:
        # one cell at a time: a sample of the cells, the whole horizon
        def run(c):
            orc.cle(net, k, dt=dt, n_steps=steps, cell_num=cells_max, cell_lo=0, cell_hi=c, seed=1, init=init,
                    threads=threads)
        cells, el = _sized(run, cells_max, 4 * threads, target_s, quantum=threads)
    cgs = float(cells) * net.G * steps
    log(f"cpu {name} {label[:12]} x{threads}: {cgs / el:.3e} c*g*s/s ({cells} cells, {el:.1f} s)")
    return {"workload": label, "restatement": name, "threads": threads, "cells": cells, "steps": steps,
            "seconds": el, "value": cgs / el, "unit": "cell*gene*steps/s"}


def cpu_baselines(target_s: float = 4.0) -> tuple:
    """CPU restatements of the reference CPU paths, timed on this host.

    cle_ref (oracle/cpu_ref.py): cle.jl:34-130's cell-vectorised structure,
      regulation as two OpenBLAS products per step, fused elementwise passes,
      two ziggurat noise matrices per step (randn!); its threads are the BLAS
      threads (the elementwise passes run on one, as Julia's broadcasts do).
    clefast_ref (oracle/, orc_cle_simulate): cle_optimized.jl:48-152's
      structure, one cell at a time with a dense G x G product, OpenMP over
      cells (CLEFast's Polyester threads), FP64.
    Each runs at 1 thread and at all threads used, on the G=5 workload (5000
    cells from zero counts, 5000 steps; full horizon, cell sample), a G=20
    telegraph slice and a G=256 slice: cle_ref runs up to 16384 cells of the
    workload at once over a sample of its steps, clefast_ref a sample of the
    cells over the whole horizon (at most 400 steps at G=256); c*g*s/s is the
    same for any such slice of a run whose per-step cost is constant.  Returns (primary, entries): the
    primary baseline is cle_ref at all threads on the G=256 slice, the
    workload of `value`."""
    from scrna_amd import samplers

    T_all = _threads_all()
    w0 = samplers.cpu_g5_workload()
    w1 = samplers.gpu_g20_workload()
    g256 = samplers.dense_g256_network()
    kin = samplers.WORKLOAD_KINETICS
    cases = [("G=5 workload (5000 cells, 5000 steps, zero init)", w0["net"], w0["kin"], w0["dt"], 5000, 5000, "zeros"),
             ("G=20 telegraph workload slice (5000 steps)", w1["net"], w1["kin"], w1["dt"], 5000, 100_000, "steady"),
             ("G=256 precision workload slice (dt=0.005)", g256, kin, 0.005, 400, 65536, "steady")]
    entries = []
    primary = None
    for case in cases:
        label, net = case[0], case[1]
        for name in ("cle_ref", "clefast_ref"):
            if name == "clefast_ref" and net.G > 64:
                continue          # CLEFast specialises on G (StaticArrays); the reference has no G=256 form of it
            for threads in (1, T_all):
                e = time_restatement(name, case, threads, target_s)
                entries.append(e)
                cells, steps, el = e["cells"], e["steps"], e["seconds"]
                cgs = float(cells) * net.G * steps
                if name == "cle_ref" and net.G == 256 and threads == T_all:
                    primary = {"value": cgs / el, "unit": "cell*gene*steps/s", "cores": threads, "kind": "port",
                               "sample": f"{cells} of 65536 cells x G=256 x {steps} of the 2000 steps of the precision "
                                         f"workload, {el:.1f} s: cle.jl's cell-vectorised FP64 form (OpenBLAS products "
                                         f"on {threads} threads, fused elementwise passes, ziggurat noise), host "
                                         f"{host_cpu()['model']}"}
    return primary, entries


def timed_runs(plan, steps: int, warmup: int):
    for _ in range(warmup):
        plan.launch()
    plan.sync()
    log(f"warmup done ({warmup}), timing {steps} runs")
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.launch()
        ms, _ = plan.stepper_ms()          # waits for this run's stepper kernels
        kernel_ms.append(ms)
    plan.sync()
    return time.perf_counter() - t0, kernel_ms


def source_digest(files) -> str:
    import hashlib
    h = hashlib.sha256()
    for f in files:
        h.update(open(os.path.join(ROOT, "syntheticscrnaseq.jl_amd", "csrc", f), "rb").read())
    return h.hexdigest()[:16]


def measured_traffic(name: str, sources, applies: bool):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (profiles/<name>, copy-calibrated).  null when the file is
    absent, was measured on other kernel sources than the ones built now, or
    the workload differs from the profiled one."""
    path = os.path.join(ROOT, "profiles", name)
    if not applies or not os.path.exists(path):
        return None
    rec = json.load(open(path))
    if rec.get("source_digest") != source_digest(sources):
        log(f"{name}: measured on other kernel sources; traffic reported as null")
        return None
    return rec["traffic_bytes"]


def run_g20(a, sc) -> dict:
    from scrna_amd import samplers

    wl = samplers.gpu_g20_workload()
    if a.cells:
        wl["cells"] = a.cells
    net, kin = wl["net"], wl["kin"]
    plan = sc.Plan([net], sc.CLE(wl["dt"]), kin, cell_num=wl["cells"], T=wl["T"], readout="mrna",
                   capture=wl["capture"], seed=20251203, init=wl["init"], snapshot_every=wl["snapshot_every"])
    steps_per_run = sc.n_steps(wl["T"], wl["dt"])
    cgs_per_run = float(net.G) * wl["cells"] * steps_per_run
    el, kernel_ms = timed_runs(plan, a.steps, a.warmup)
    res = plan.download(0)
    if not np.all(np.isfinite(res.y)):
        raise SystemExit("non-finite state")
    value = cgs_per_run * a.steps / el
    k_ms = float(np.mean(kernel_ms))
    census = valu_census(net)
    achieved = census["fp32_equiv_flops"] * cgs_per_run / (k_ms * 1e-3) / 1e12
    traffic = measured_traffic("cle_stepper_g20_traffic.json", ("cle_kernels.hip", "hill.h", "device_rng.h", "blob.h"),
                               wl["cells"] == 100_000)
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s (FP32-FMA-equivalent VALU issue)",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "cle_stepper_kernel", "kernel_ms": k_ms,
                     "census_per_cgs": census},
    }
    if a.tile:
        out["config"]["genes_per_lane"] = a.tile
    if not a.no_cpu and a.workload == "g20":
        log("cpu baseline")
        e = time_restatement("clefast_ref", ("G=20 telegraph workload", net, kin, wl["dt"], steps_per_run,
                                             wl["cells"], wl["init"]), _threads_all(), 10.0)
        out["cpu_baseline"] = {"value": e["value"], "unit": e["unit"], "cores": e["threads"], "kind": "port",
                               "sample": f"{e['cells']} of {wl['cells']} cells x G=20 x {steps_per_run} steps (full "
                                         f"horizon), {e['seconds']:.1f} s: the FP64 per-cell restatement (CLEFast's "
                                         f"structure) with telegraph + AND/OR, OpenMP over cells, host "
                                         f"{host_cpu()['model']}"}
    plan.close()
    return out


def run_g256(a, sc) -> dict:
    """Three contraction forms at G=256 (65,536 cells, 2000 steps, dt=0.005) and
    their per-gene mean/variance error against an FP64 replay of a cell subset
    with the identical Philox stream."""
    from scrna_amd import samplers

    net = samplers.dense_g256_network()
    kin = samplers.WORKLOAD_KINETICS
    cells = a.cells or 65536
    dt, steps = 0.005, 2000
    T = steps * dt
    n_sub = a.subset
    out = {"workload": "G=256 additive network, 25% dense A, 8 AND + 8 OR, 65536 cells, 2000 steps dt=0.005",
           "cells": cells, "steps": steps, "results": {}}
    subsets = {}
    precs = [p for p in ("fp32_unfused", "fp32_valu", "fp32_mfma", "bf16", "bf16x3") if p in a.precisions.split(",")]
    for prec in precs:
        plan = sc.Plan([net], sc.CLE(dt), kin, cell_num=cells, T=T, readout="both", seed=777, precision=prec)
        # the timed forms of `value` run the requested count; comparison forms at most 5 (they take up to 1.3 s each)
        n_runs = a.steps if prec == a.headline_precision else min(a.steps, 5)
        el, kms = timed_runs(plan, n_runs, min(a.warmup, 2))
        cgs = float(net.G) * cells * steps
        k_ms = float(np.mean(kms))
        census = valu_census(net, dense=prec in ("bf16", "bf16x3", "fp32_mfma"))
        r = {"cgs_per_s": cgs * n_runs / el, "kernel_ms": k_ms, "runs": n_runs, "ms_per_run": el * 1e3 / n_runs,
             "valu_frac": census["fp32_equiv_flops"] * cgs / (k_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
             "census_per_cgs": census}
        if prec in ("bf16", "bf16x3", "fp32_mfma"):
            mfma_flops = 2.0 * net.G * cgs * (3 if prec == "bf16x3" else 1)
            r["mfma_tflops_issued"] = mfma_flops / (k_ms * 1e-3) / 1e12
            if prec == "fp32_mfma":
                # bound: the contraction at the FP32 matrix peak, or the elementwise VALU census
                t_bound = max(2.0 * net.G * cgs / (FP32_PEAK_TFLOPS * 1e12),
                              census["fp32_equiv_flops"] * cgs / (FP32_PEAK_TFLOPS * 1e12))
                r["roofline_frac"] = t_bound / (k_ms * 1e-3)
        y = plan.download(0).y
        subsets[prec] = y[:n_sub]
        out["results"][prec] = r
        log(f"{prec}: {r['cgs_per_s']:.3e} c*g*s/s, kernel {k_ms:.1f} ms")
        plan.close()
    if n_sub > 0 and not a.no_cpu:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        ref = orc.cle(net, kin, dt=dt, n_steps=steps, cell_num=cells, cell_lo=0, cell_hi=n_sub, seed=777,
                      threads=max(1, min(16, os.cpu_count() or 1)))
        log(f"FP64 replay of {n_sub} cells: {time.perf_counter() - t0:.1f} s")
        G = net.G
        for prec, y in subsets.items():
            errs = {}
            for k, want in ((0, ref["m"]), (1, ref["p"])):
                got = y[:, k * G:(k + 1) * G].T
                me = np.abs(got.mean(1) - want.mean(1)) / np.maximum(np.abs(want.mean(1)), 1e-3)
                ve = np.abs(got.var(1) - want.var(1)) / np.maximum(want.var(1), 1e-3)
                errs[("mrna", "protein")[k]] = {"max_mean_rel_err": float(me.max()), "max_var_rel_err": float(ve.max())}
            out["results"][prec]["vs_fp64_replay"] = errs
        out["replay_cells"] = n_sub
    res = out["results"]
    fused = [p for p in ("fp32_mfma", "fp32_valu") if p in res]
    best = max(fused, key=lambda p: res[p]["cgs_per_s"]) if fused else None
    out["fp32_fused_form"] = {"fp32_mfma": "FP32 dense contraction on the matrix cores (v_mfma_f32_16x16x4_f32), "
                                           "FP32 products and accumulation, fused stepper",
                              "fp32_valu": "sparse FP32 VALU stepper, 16 genes per lane, in-edges streamed per step",
                              None: None}[best]
    for prec in ("fp32_valu", "fp32_mfma", "bf16", "bf16x3"):
        if prec not in res:
            continue
        if best:
            res[prec]["speedup_vs_fp32_fused"] = res[prec]["cgs_per_s"] / res[best]["cgs_per_s"]
        if "fp32_unfused" in res:
            res[prec]["speedup_vs_fp32_reference_structured"] = res[prec]["cgs_per_s"] / res["fp32_unfused"]["cgs_per_s"]
    return out


def run_wide(a, sc) -> dict:
    """Networks above the fused steppers' 256 genes: the wide path (one library
    GEMM and one fused elementwise kernel per step, FP32 / BF16 / BF16x3 products)
    against the reference-structured FP32 path, on the G=256 workload's network
    construction at --genes genes (65,536 cells, 2000 steps, dt=0.005)."""
    from scrna_amd import samplers

    G = a.genes
    net = samplers.dense_network(G)
    kin = samplers.WORKLOAD_KINETICS
    cells = a.cells or 65536
    dt, steps = 0.005, 2000
    T = steps * dt
    cgs = float(G) * cells * steps
    out = {"workload": f"G={G} additive network, 25% dense A, 8 AND + 8 OR, {cells} cells, {steps} steps dt=0.005",
           "cells": cells, "steps": steps, "results": {}}
    subsets = {}
    n_sub = min(a.subset, 32)            # the FP64 replay costs G^2/4 edge updates per cell and step
    precs = []
    for p in a.precisions.split(","):
        p = "fp32" if p == "fp32_mfma" else p
        if p in ("fp32_unfused", "fp32", "bf16", "bf16x3") and p not in precs:
            precs.append(p)
    for prec in precs:
        plan = sc.Plan([net], sc.CLE(dt), kin, cell_num=cells, T=T, readout="both", seed=777, precision=prec)
        n_runs = min(a.steps, 3)
        el, kms = timed_runs(plan, n_runs, min(a.warmup, 1))
        k_ms = float(np.mean(kms))
        r = {"cgs_per_s": cgs * n_runs / el, "kernel_ms": k_ms, "runs": n_runs, "ms_per_run": el * 1e3 / n_runs,
             "gemm_tflops": 2.0 * G * cgs * (3 if prec == "bf16x3" else (2 if prec == "fp32_unfused" else 1))
             / (k_ms * 1e-3) / 1e12}
        if n_sub > 0 and not a.no_cpu:
            subsets[prec] = plan.download(0).y[:n_sub]
        out["results"][prec] = r
        log(f"{prec}: {r['cgs_per_s']:.3e} c*g*s/s, {k_ms:.1f} ms per run")
        plan.close()
    if subsets and not a.no_cpu:
        from oracle import oracle as orc
        ref = orc.cle(net, kin, dt=dt, n_steps=steps, cell_num=cells, cell_lo=0, cell_hi=n_sub, seed=777,
                      threads=max(1, min(16, os.cpu_count() or 1)))
        for prec, y in subsets.items():
            errs = {}
            for k, want in ((0, ref["m"]), (1, ref["p"])):
                got = y[:, k * G:(k + 1) * G].T
                me = np.abs(got.mean(1) - want.mean(1)) / np.maximum(np.abs(want.mean(1)), 1e-3)
                ve = np.abs(got.var(1) - want.var(1)) / np.maximum(want.var(1), 1e-3)
                errs[("mrna", "protein")[k]] = {"max_mean_rel_err": float(me.max()), "max_var_rel_err": float(ve.max())}
            out["results"][prec]["vs_fp64_replay"] = errs
        out["replay_cells"] = n_sub
    res = out["results"]
    for prec in ("fp32", "bf16", "bf16x3"):
        if prec in res and "fp32_unfused" in res:
            res[prec]["speedup_vs_fp32_reference_structured"] = res[prec]["cgs_per_s"] / res["fp32_unfused"]["cgs_per_s"]
        if prec in res and "fp32" in res:
            res[prec]["speedup_vs_fp32_wide"] = res[prec]["cgs_per_s"] / res["fp32"]["cgs_per_s"]
    return out


def run_headline(a, sc) -> dict:
    """The metric line: value = BF16-MFMA throughput at G=256 with the precision
    comparison, the G=20 workload and the serving workload as fields."""
    from scrna_amd import samplers

    a.headline_precision = "bf16"
    g = run_g256(a, sc)
    res = g["results"]
    bf = res["bf16"]
    net = samplers.dense_g256_network()
    cgs = float(net.G) * g["cells"] * g["steps"]
    census = bf["census_per_cgs"]
    achieved = census["fp32_equiv_flops"] * cgs / (bf["kernel_ms"] * 1e-3) / 1e12
    traffic = measured_traffic("cle_dense_g256_traffic.json", ("dense_kernels.hip", "device_rng.h", "blob.h"),
                               g["cells"] == 65536)
    errs = bf.get("vs_fp64_replay", {})
    worst = max([v for sp in errs.values() for v in sp.values()] or [float("nan")])
    log("g20 workload")
    g20 = run_g20(a, sc)
    log("serving workload")
    srv = run_serve(a, sc)
    log("tau-leap workload")
    tl_args = argparse.Namespace(**{**vars(a), "steps": min(a.steps, 3), "warmup": 1, "no_cpu": True})
    tl = run_tauleap(tl_args, sc)
    wide = None
    if not a.no_wide:
        log("wide workload (G=1024)")
        w_args = argparse.Namespace(**{**vars(a), "steps": 1, "warmup": 1, "no_cpu": True, "genes": 1024,
                                       "precisions": "fp32_unfused,fp32,bf16"})
        try:
            wr = run_wide(w_args, sc)
            wide = {"workload": wr["workload"],
                    **{p: {k: r[k] for k in ("cgs_per_s", "kernel_ms") if k in r} for p, r in wr["results"].items()},
                    "bf16_vs_fp32_reference_structured":
                        wr["results"]["bf16"].get("speedup_vs_fp32_reference_structured"),
                    "bf16_vs_fp32_wide": wr["results"]["bf16"].get("speedup_vs_fp32_wide")}
        except Exception as e:      # the headline stands without this field
            wide = {"error": f"{type(e).__name__}: {e}"}
    out = {
        "metric": "cell*gene*steps/s, BF16-MFMA vs FP32 CLE at <=1% moment err; p99 latency",
        "value": bf["cgs_per_s"],
        "unit": "cell*gene*steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": bf["ms_per_run"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded G=256 network, Philox noise)",
        "config": {"workload": "G=256 additive network, 25% dense A (+-U(0.5,3)), 8 AND + 8 OR, K=10 n=2, constitutive, "
                               "65536 cells, 2000 steps dt=0.005; regulation contraction in BF16 on MFMA with FP32 "
                               "accumulation, FP32 state",
                   "G": net.G, "cells": g["cells"], "sim_steps": g["steps"], "parallelism": "single GPU"},
        "headline": {
            "bf16_vs_fp32_fused": bf.get("speedup_vs_fp32_fused"),
            "bf16_vs_fp32_reference_structured": bf.get("speedup_vs_fp32_reference_structured"),
            "bf16_max_moment_rel_err_vs_fp64": worst,
            "bf16_within_1pct": bool(worst <= 0.01),
            "fp32_fused_form": g.get("fp32_fused_form"),
            "serving_qps": srv["qps"], "serving_p50_ms": srv["p50_ms"], "serving_p99_ms": srv["p99_ms"],
            "serving_p99_bound_ms": srv["p99_bound_ms"], "serving_p99_within_bound": srv["p99_within_bound"],
        },
        "precision": g,
        "g20": {k: g20[k] for k in ("value", "ms_per_step", "config", "roofline")},
        "serving": srv,
        "tauleap": {k: tl.get(k) for k in ("workload", "cgs_per_s", "kernel_ms", "roofline")},
        "wide_g1024": wide,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s (FP32-FMA-equivalent VALU issue)", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic, "kernel": "cle_dense_kernel (BF16 MFMA)", "kernel_ms": bf["kernel_ms"],
                     "census_per_cgs": census,
                     "mfma_tflops": bf.get("mfma_tflops_issued"),
                     "mfma_frac": (bf.get("mfma_tflops_issued") or 0.0) / BF16_PEAK_TFLOPS},
    }
    if not a.no_cpu:
        log("cpu baselines")
        primary, entries = cpu_baselines()
        out["cpu_baseline"] = primary
        out["cpu_baselines"] = {"host": host_cpu(), "threads_all": _threads_all(), "runs": entries}
    return out


def tauleap_census():
    """Executed VALU work per cell*gene*step of the tau-leap stepper on the G=20
    workload, priced in FP32-FMA issue units.  Its samplers (inversion and
    rejection, tauleap_binomial.jl) do data-dependent work, so there is no fixed
    operation count: the census is the kernel's own instruction mix, counted by
    rocprofv3 (profiles/tauleap_stepper_g20_mix.json, valid while the kernel
    sources match its digest).  It includes the compiler's moves and address
    arithmetic, so its fraction of peak is an issue-efficiency figure, an upper
    bound on the algorithmic one."""
    path = os.path.join(ROOT, "profiles", "tauleap_stepper_g20_mix.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    if rec.get("source_digest") != source_digest(rec["sources"]):
        log("tauleap_stepper_g20_mix.json: measured on other kernel sources; census reported as null")
        return None
    m = rec["lane_instructions_per_cgs"]
    classes = {"fma_f64": (m["SQ_INSTS_VALU_FMA_F64"], COST_F64_FMA), "add_f64": (m["SQ_INSTS_VALU_ADD_F64"], COST_F64_ADD),
               "mul_f64": (m["SQ_INSTS_VALU_MUL_F64"], COST_F64_MUL),
               "trans_f64": (m["SQ_INSTS_VALU_TRANS_F64"], COST_F64_TRANS),
               "int64": (m["SQ_INSTS_VALU_INT64"], COST_MUL32), "int32": (m["SQ_INSTS_VALU_INT32"], COST_INT),
               "cvt": (m["SQ_INSTS_VALU_CVT"], COST_CVT)}
    other = m["SQ_INSTS_VALU"] - sum(n for n, _ in classes.values())
    classes["other"] = (other, COST_OTHER)
    fma_eq = sum(n * c for n, c in classes.values())
    return {"lane_instructions": {k: round(n, 3) for k, (n, _) in classes.items()},
            "valu_instructions": m["SQ_INSTS_VALU"], "fma_equiv_ops": fma_eq, "fp32_equiv_flops": 2.0 * fma_eq}


def run_tauleap(a, sc) -> dict:
    """GPU binomial tau-leap on the G=20 benchmark network: 100,000 cells, 5000
    steps (dt=0.01), exact Poisson / binomial draws, FP64 propensities."""
    from scrna_amd import samplers

    wl = samplers.gpu_g20_workload()
    cells = a.cells or wl["cells"]
    net, kin = wl["net"], wl["kin"]
    dt, T = wl["dt"], wl["T"]
    plan = sc.Plan([net], sc.BinomialTauLeap(dt), kin, cell_num=cells, T=T, readout="both", seed=31)
    steps = sc.n_steps(T, dt)
    cgs = float(net.G) * cells * steps
    el, kms = timed_runs(plan, a.steps, a.warmup)
    y = plan.download(0).y
    k_ms = float(np.mean(kms))
    out = {"workload": f"binomial tau-leap, G=20 telegraph network, {cells} cells, {steps} steps dt={dt}",
           "cgs_per_s": cgs * a.steps / el, "kernel_ms": k_ms, "dtype": "f64 propensities, i32 counts",
           "mean_mrna": float(y[:, :net.G].mean()), "mean_protein": float(y[:, net.G:].mean())}
    census = tauleap_census() if cells == wl["cells"] else None
    if census:
        achieved = census["fp32_equiv_flops"] * cgs / (k_ms * 1e-3) / 1e12
        out["roofline"] = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                           "unit": "TFLOP/s (FP32-FMA-equivalent VALU issue, executed instruction mix)",
                           "frac": achieved / FP32_PEAK_TFLOPS, "kernel": "tauleap_stepper_kernel",
                           "census_per_cgs": census}
    if not a.no_cpu:
        from oracle import oracle as orc
        threads = max(1, min(16, os.cpu_count() or 1))
        probe = 16 * threads
        t0 = time.perf_counter()
        orc.tauleap(net, kin, dt=dt, n_steps=steps, cell_num=cells, cell_lo=0, cell_hi=probe, seed=31,
                    threads=threads)
        rate = probe / max(time.perf_counter() - t0, 1e-6)
        sample = int(min(cells, max(probe, rate * 12.0)))
        sample = max(threads, (sample // threads) * threads)
        t0 = time.perf_counter()
        orc.tauleap(net, kin, dt=dt, n_steps=steps, cell_num=cells, cell_lo=0, cell_hi=sample, seed=31,
                    threads=threads)
        el_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": float(net.G) * sample * steps / el_cpu, "unit": "cell*gene*steps/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{sample} of {cells} cells, full horizon, FP64 C restatement of "
                                         f"tauleap_binomial.jl, OpenMP over cells, {el_cpu:.1f} s"}
    plan.close()
    return out


def run_sbi(a, sc) -> dict:
    """Simulation-based-inference batch: draw networks from the native prior
    (kinetics.jl:105-205), simulate them in one plan, capture counts."""
    n, G, cells = a.networks if a.networks != 4096 else 1024, 20, a.cells or 1024
    dt, T = 0.01, a.T or 50.0
    t0 = time.perf_counter()
    nets = sc.prior.sample_networks(G, n, seed=7)
    t_prior = time.perf_counter() - t0
    kin = sc.KineticParams(k_t=5.0, K_d=10.0, n=2.0, mu_m=1.0, mu_p=0.2)
    cap = sc.CaptureModel(efficiency=0.1, efficiency_std=0.3, readout="mrna")
    t0 = time.perf_counter()
    plan = sc.Plan(nets, sc.CLE(dt), kin, cell_num=cells, T=T, readout="mrna", capture=cap, seed=11)
    t_build = time.perf_counter() - t0
    el, kms = timed_runs(plan, a.steps, a.warmup)
    steps = sc.n_steps(T, dt)
    cgs = float(n) * G * cells * steps
    plan.close()
    return {"workload": f"SBI batch: {n} prior networks (G={G}), {cells} cells each, {steps} steps, capture",
            "networks": n, "prior_s": t_prior, "plan_build_s": t_build, "cgs_per_s": cgs * a.steps / el,
            "kernel_ms": float(np.mean(kms)), "networks_per_s": n * a.steps / el}


def run_batch(a, sc) -> dict:
    """4096 uneven networks with Moran population dynamics in one hipGraph."""
    from scrna_amd import samplers

    n = a.networks
    nets, cells, pop = samplers.batch_networks(n)
    if a.div_interval:
        import dataclasses
        pop = dataclasses.replace(pop, div_check_interval=a.div_interval)
    kin = samplers.WORKLOAD_KINETICS
    T, dt = a.T or 100.0, 0.01
    # Tile shapes (genes per lane per G bucket) from the autotuner, timed on this device.
    t0 = time.perf_counter()
    tune_path = os.path.join(tempfile.mkdtemp(), "tiles.json")
    sc.default_context().autotune(max_G=max(x.G for x in nets), cache_path=tune_path,
                                  phase_steps=pop.div_check_interval)
    t_tune = time.perf_counter() - t0
    tile_table = json.load(open(tune_path))["table"]
    pops = [sc.PopulationConfig(cell_num=c, growth_rate=pop.growth_rate, V_div=pop.V_div, V_init=pop.V_init,
                                div_check_interval=pop.div_check_interval, partition=pop.partition,
                                replace=pop.replace) for c in cells]
    t0 = time.perf_counter()
    if a.no_population:   # the same networks and cells as independent cells, one horizon launch
        plan = sc.Plan(nets, sc.CLE(dt), kin, cell_num=list(cells), T=T, readout="mrna", seed=99)
    else:
        plan = sc.Plan(nets, sc.CLE(dt), kin, population=pops, T=T, readout="mrna", seed=99)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    if not a.no_graph:
        plan.use_graph()
    t_graph = time.perf_counter() - t0
    el, kms = timed_runs(plan, a.steps, a.warmup)
    steps = sc.n_steps(T, dt)
    cgs = float(sum(net.G * c for net, c in zip(nets, cells))) * steps
    sample = plan.download(0)
    popdesc = ("no population" if a.no_population else
               f"population (growth 0.05, V_div 2, binomial partition, Moran, check every {pop.div_check_interval})")
    return {"workload": f"{n} networks, G~U{{5..64}}, cells logU[256,8192], {popdesc}, telegraph, T={T}, dt={dt}, "
                        f"{'direct launches' if a.no_graph else 'one hipGraph'}",
            "networks": n, "total_cells": int(sum(cells)), "mean_G": float(np.mean([x.G for x in nets])),
            "cell_gene_steps": cgs, "cgs_per_s": cgs * a.steps / el, "s_per_run": el / a.steps,
            "graph_kernel_ms": float(np.mean(kms)), "plan_build_s": t_build, "graph_capture_s": t_graph,
            "autotune_s": t_tune, "tile_table": tile_table, "finite": bool(np.all(np.isfinite(sample.y)))}


# Stated p99 latency bound of the serving workload at 4 requests/s: twice the
# time the largest request this stream can draw (G=200, 50,000 cells, 20,000
# steps, 2e11 cell*gene*steps) takes alone at the G=200 stepper's rate.
SERVE_P99_BOUND_MS = 1000.0


def run_serve(a, sc) -> dict:
    """Poisson arrivals at a fixed QPS; G in {10,50,200}, cells logU[1k,50k],
    T ~ U(10,200) at dt=0.01, half telegraph, capture on every request."""
    from scrna_amd import samplers

    rng = np.random.default_rng(5)
    kin = samplers.WORKLOAD_KINETICS
    qps, n_req = a.qps, a.requests
    reqs = []
    for i in range(n_req):
        G = int(rng.choice([10, 50, 200]))
        tele = bool(i % 2)
        net = samplers.mixed_network(G, rng, n_tf=(2, 3) if G <= 10 else (4, 5),
                                     targets_per_tf=(2, 4) if G <= 10 else (3, 6), n_and=(0, 1) if G <= 10 else (1, 2),
                                     n_or=(0, 1) if G <= 10 else (1, 2), telegraph=tele)
        cells = int(np.exp(rng.uniform(np.log(1000), np.log(50000))))
        T = float(rng.uniform(10, 200))
        reqs.append((net, cells, T))
    gaps = rng.exponential(1.0 / qps, size=n_req)
    cap = sc.CaptureModel(0.1, 0.3, "mrna")
    with sc.Server(max_slots=64) as srv:
        t_start = time.perf_counter()
        tickets = []
        t_next = t_start
        for i, (net, cells, T) in enumerate(reqs):
            t_next += gaps[i]
            while time.perf_counter() < t_next:
                time.sleep(min(0.01, t_next - time.perf_counter()))
            tickets.append(srv.submit(net, sc.CLE(0.01), kin, cell_num=cells, T=T, readout="mrna", capture=cap,
                                      seed=1000 + i))
            if i % 10 == 9:
                log(f"submitted {i + 1}/{n_req}")
        for t in tickets:
            srv.wait(t)
        wall = time.perf_counter() - t_start
        st = srv.stats()
    return {"workload": f"Poisson arrivals at {qps} req/s, {n_req} requests, G in {{10,50,200}}, cells logU[1k,50k], "
                        "T~U(10,200) dt=0.01, half telegraph, LogNormal-Binomial capture",
            "qps": qps, "requests": n_req, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
            "p99_bound_ms": SERVE_P99_BOUND_MS, "p99_within_bound": bool(st["p99_ms"] <= SERVE_P99_BOUND_MS),
            "mean_ms": st["mean_ms"],
            "max_ms": st["max_ms"], "completed": st["completed"], "wall_s": wall,
            "aggregate_cgs_per_s": st["cell_gene_steps"] / wall, "busy_cgs_per_s": st["cell_gene_steps"] / max(st["busy_s"], 1e-9)}


def run_copy(a) -> dict:
    """Plain device-to-device copy of known size (hipMemcpy's blit kernel), used
    to scale the FETCH_SIZE / WRITE_SIZE hardware counters to bytes."""
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = a.bytes
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0     # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    return {"workload": "device copy", "bytes": nbytes, "copies": 3}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["headline", "g20", "g256", "batch", "serve", "copy", "tauleap", "sbi",
                                           "wide"],
                    default="headline")
    ap.add_argument("--cells", type=int, default=None, help="override cell count (testing only)")
    ap.add_argument("--tile", type=int, default=0, help="genes per lane for the FP32 stepper (2, 4, 8, 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--subset", type=int, default=1024, help="g256: cells replayed by the FP64 oracle")
    ap.add_argument("--precisions", default="fp32_unfused,fp32_valu,fp32_mfma,bf16,bf16x3",
                    help="g256: comma-separated forms")
    ap.add_argument("--networks", type=int, default=4096, help="batch: number of networks")
    ap.add_argument("--T", type=float, default=None, help="batch: horizon")
    ap.add_argument("--no-graph", action="store_true", help="batch: launch the plan's kernels directly, no hipGraph")
    ap.add_argument("--div-interval", type=int, default=0, help="batch: division check interval (default 10)")
    ap.add_argument("--no-population", action="store_true",
                    help="batch: the same cells without population dynamics, one horizon launch")
    ap.add_argument("--qps", type=float, default=4.0, help="serve: arrival rate")
    ap.add_argument("--requests", type=int, default=60, help="serve: number of requests")
    ap.add_argument("--bytes", type=int, default=1 << 30, help="copy: bytes per copy")
    ap.add_argument("--genes", type=int, default=1024, help="wide: genes of the network (above 256)")
    ap.add_argument("--no-wide", action="store_true", help="headline: skip the G=1024 wide-path field")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(cells and networks would shard across GPUs with no exchange during a run).",
              file=sys.stderr)
        return 2
    if a.workload == "copy":
        print(json.dumps(run_copy(a)))
        return 0

    import scrna_amd as sc

    if a.tile:
        sc.default_context().set_tile(1, min(256, 16 * a.tile), a.tile)
    a.headline_precision = None
    runner = {"headline": run_headline, "g20": run_g20, "g256": run_g256, "batch": run_batch, "serve": run_serve,
              "tauleap": run_tauleap, "sbi": run_sbi, "wide": run_wide}[a.workload]
    print(json.dumps(runner(a, sc)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
