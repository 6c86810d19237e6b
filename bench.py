#!/usr/bin/env python3
"""PairHMM GCUPS on one MI355X: batched uneven active regions in one persistent launch.

A step is one forward pass of the whole batch (planning, FP32 anti-diagonal
kernel, FP64 rescue) over inputs already resident in HBM.  The default
workload is 4096 uneven regions (2-32 haplotypes of 150-600 bp, 2-30 or
200-2000 reads of 50-250 bp, bimodal Phred), seed 41002.

The JSON line carries:
* value      -- whole-job GCUPS = sum(n*m) * steps / wall seconds
* roofline   -- the FP32 kernels' rate (11 flops per computed cell) over their
                event-timed span (HIP events before the first class kernel
                and after the last; consecutive classes overlap on two
                streams), against MI355X's FP32 vector peak;
                `effective` counts the pairs' n*m instead (haplotype prefix
                sharing computes fewer cells than that)
* cpu_baseline -- the reference's CPU path (trie_many + pairhmm_pairs
                fallback, compiled from the reference sources into oracle/_ref)
                on a time-bounded sample of the same workload, on host threads
* graph      -- the same step replayed from a captured HIP graph
* plain      -- the one-pair-per-launch GPU forward on a sample of the workload
* ref_design -- the reference's own GPU design (one thread per pair, DP rows in
                global memory, FP32 with its 2^-40 rescale), batched in length
                buckets, on a sample of whole regions of the same workload
* fwd_bwd    -- forward + adjoint (gradients wrt gap, ext, 64-bin error table)
                on the 1024-region HMM-sampled workload (seed 41004, 2.05M pairs);
                `fused_*` counts the gradient call alone, which also returns
                ln L per pair
* genotype   -- the chromosome-scale sample (seed 41005): forward then the
                diploid genotype likelihood over every haplotype pair of every
                400 bp region (normalized rows, certainty filter, molecule rows,
                allele-balance + artifact mixture), end to end on the device,
                at 30x and 2x depth; the CPU chain (C restatement of the
                reference forward + genotype) timed on a bounded sample

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--regions N]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "PairHMM GCUPS (fwd, fwd+bwd) at <=1e-4 log10 err vs FP64; % of FP32 peak"
FP32_PEAK_TFLOPS = 157.3
FLOPS_PER_CELL = 11
CPU_THREAD_CAP = 16   # host threads per GPU on a shared multi-GPU node
REFERENCE_GAP_EXT = (0.001, 0.1)   # the model trie_many / pairhmm_pairs compute (batch_hmm.cpp:30)


def available_cpus() -> int:
    """Affinity intersected with the cgroup CPU quota (cpu_schedule.py:5-18)."""
    count = len(os.sched_getaffinity(0))
    try:
        for line in Path("/proc/self/cgroup").read_text().splitlines():
            if line.startswith("0::"):
                path = Path("/sys/fs/cgroup") / line.split(":", 2)[2].lstrip("/")
                for d in [path, *path.parents]:
                    f = d / "cpu.max"
                    if f.exists():
                        quota, period = f.read_text().split()
                        if quota != "max":
                            count = min(count, max(1, math.floor(int(quota) / int(period))))
                        break
    except (OSError, ValueError):
        pass
    return max(1, count)


def workload(name: str, regions: int):
    from vardepth_amd import workloads
    if name == "uneven":
        return workloads.uneven_regions(seed=41002, regions=regions), "uneven-regions (4096 active regions, seed 41002)"
    if name == "single":
        return workloads.single_region(seed=41001), "single-region 64x8 (seed 41001)"
    if name == "stress":
        return workloads.precision_stress(seed=41003, regions=regions), "precision-stress (seed 41003)"
    raise ValueError(name)


def measured_traffic(args):
    """HBM bytes per k_fwd32 launch from the committed PMC summary (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated) when it was measured on this workload, else None."""
    f = ROOT / "profiles" / "k_fwd32_uneven_regions_4096_traffic.json"
    if args.workload != "uneven" or args.regions != 4096 or args.mode != 0 or args.precision != "mixed":
        return None
    try:
        t = json.loads(f.read_text())
        return int(t["fetch_bytes"] + t["write_bytes"])
    except (OSError, ValueError, KeyError):
        return None


def fp32_cells(batch, mode: int) -> int:
    """Cells of the pairs the FP32 kernel processes (reads of <= 256 rows)."""
    rl = batch.read_lengths() + (1 if mode == 1 else 0)
    hl = batch.hap_lengths()
    total = 0
    for k in range(batch.nregions):
        r = rl[batch.region_reads[k]:batch.region_reads[k + 1]]
        r = r[(r > 0) & (r <= 256)] - (1 if mode == 1 else 0)
        total += int(r.sum()) * int(hl[batch.region_haps[k]:batch.region_haps[k + 1]].sum())
    return total


def reference_region(ref, batch, k, mode):
    """The reference CPU chain on region k (trie_vector.py:16-42): trie_many over
    the reads sorted by length and the sorted haplotypes, NaN values recomputed
    by pairhmm_pairs; returned in the batch's layout (row-major reads x haplotypes
    in batch order).  Those two hard-code g = .001, x = .1 (batch_hmm.cpp:30,
    trie_vector.cpp:219); with other model parameters, or with ref None, the
    oracle's C restatement of that chain (parametric) is used."""
    import oracle
    if ref is None or (batch.gap, batch.ext) != REFERENCE_GAP_EXT:
        return oracle.regions_cpu(batch.subset([k]), mode=mode, threads=1)[0]
    reads, quals, haps = batch.region(k)
    ri = sorted(range(len(reads)), key=lambda i: len(reads[i]))   # trie_vector.py:30
    hi = sorted(range(len(haps)), key=lambda j: haps[j])          # trie_vector.py:27
    out, _ = ref.trie_many([reads[i] for i in ri], [quals[i] for i in ri], [haps[j] for j in hi], mode)
    nan = np.argwhere(np.isnan(out))
    if len(nan):
        jobs = [(reads[ri[a]], quals[ri[a]], haps[hi[b]]) for a, b in nan]
        out[nan[:, 0], nan[:, 1]] = ref.pairhmm_pairs(jobs, mode)   # trie_vector.py:40-41
    block = np.empty((len(reads), len(haps)))
    block[np.ix_(ri, hi)] = out
    return block.ravel()


def cpu_baseline(batch, mode: int, seconds: float, threads: int, device_out=None):
    """Reference CPU chain on a time-bounded random sample of regions.

    With device_out (the headline step's output, batch layout), the sampled
    regions' reference values are kept and the step's per-pair log10 error
    against them is reported."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    ref = oracle.ref()
    kind = "reference" if ref is not None else "port"
    rng = np.random.default_rng(7)
    order = rng.permutation(batch.nregions)
    rl, hl = batch.read_lengths(), batch.hap_lengths()
    kept = {}

    def one(k, keep):
        vals = reference_region(ref, batch, k, mode)
        if keep:
            kept[k] = vals
        a = int(rl[batch.region_reads[k]:batch.region_reads[k + 1]].sum())
        b = int(hl[batch.region_haps[k]:batch.region_haps[k + 1]].sum())
        return a * b

    def run(nthreads, budget, keep):
        cells = 0
        done = 0
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=nthreads) as ex:
            futs = []
            it = iter(order)
            for _ in range(nthreads * 2):
                k = next(it, None)
                if k is not None:
                    futs.append(ex.submit(one, int(k), keep))
            while futs:
                f = futs.pop(0)
                cells += f.result()
                done += 1
                if time.perf_counter() - t0 < budget:
                    k = next(it, None)
                    if k is not None:
                        futs.append(ex.submit(one, int(k), keep))
        return cells, done, time.perf_counter() - t0

    cells, done, dt = run(threads, seconds, device_out is not None)
    c1, d1, t1 = run(1, max(2.0, seconds / 5), False)
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    res = {"value": cells / dt / 1e9, "unit": "GCUPS", "cores": threads, "kind": kind,
           "sample": "%d of %d regions of the same workload (%.3g cells, %.1f s wall), FP64 %s; "
                     "1 core: %.3f GCUPS over %d regions" % (
                         done, batch.nregions, cells, dt,
                         "trie_many+pairhmm_pairs compiled from the reference sources" if kind == "reference"
                         else "C restatement of trie_many+pairhmm_pairs", c1 / t1 / 1e9, d1),
           "gcups_1core": c1 / t1 / 1e9, "cpu_model": model,
           "cores_available": available_cpus(), "cores_online": os.cpu_count(),
           "core_note": "threads = min(available cores, %d): one GPU's share of the host" % CPU_THREAD_CAP}
    if device_out is not None and kept:
        res["accuracy"] = compare_regions(batch, device_out, kept)
    return res


def compare_regions(batch, device_out, expected: dict):
    """Per-pair |log10 L_device - log10 L_reference| over the regions in
    `expected` (region -> reference values in batch layout), with the exact
    agreement of NaN and -inf positions."""
    oo = batch.out_off()
    ks = sorted(expected)
    dev = np.concatenate([device_out[oo[k]:oo[k + 1]] for k in ks])
    ref = np.concatenate([expected[k] for k in ks])
    ln10 = math.log(10.0)
    nan_eq = bool(np.array_equal(np.isnan(dev), np.isnan(ref)))
    inf_eq = bool(np.array_equal(np.isneginf(dev), np.isneginf(ref)))
    fin = np.isfinite(ref) & np.isfinite(dev)
    err = np.abs(dev[fin] - ref[fin]) / ln10
    return {"regions": len(ks), "pairs": int(len(ref)),
            "max_log10_err": float(err.max()) if err.size else 0.0,
            "p99_log10_err": float(np.quantile(err, 0.99)) if err.size else 0.0,
            "within_1e-4_log10": bool(err.size == 0 or err.max() <= 1e-4),
            "nan_positions_equal": nan_eq, "neginf_positions_equal": inf_eq,
            "nonfinite_pairs": int(len(ref) - fin.sum())}


def plain_baseline(batch, mode: int, pairs: int):
    """The one-pair-per-launch forward on the first `pairs` pairs of a seeded
    random selection of regions (region order inside the selection)."""
    from vardepth_amd import engine
    rng = np.random.default_rng(11)
    order = rng.permutation(batch.nregions)
    counts = np.diff(batch.region_reads) * np.diff(batch.region_haps)
    take, total = [], 0
    for k in order:
        take.append(int(k))
        total += int(counts[k])
        if total >= pairs:
            break
    sub = batch.subset(take)
    out, ms = engine.forward_plain(sub, mode=mode, max_pairs=pairs)
    rl, hl = sub.read_lengths(), sub.hap_lengths()
    done = cells = 0
    for k in range(sub.nregions):
        for r in range(sub.region_reads[k], sub.region_reads[k + 1]):
            for h in range(sub.region_haps[k], sub.region_haps[k + 1]):
                if done < pairs:
                    cells += int(rl[r]) * int(hl[h])
                    done += 1
    return {"gcups": round(cells / ms / 1e6, 4), "pairs": done, "cells": cells, "gpu_ms": round(ms, 3),
            "kind": "one kernel launch per pair, one thread per read row, FP32, inputs resident in HBM"}


def refdesign_baseline(batch, mode: int, cells_target: float, engine_gcups: float):
    """The reference's own GPU design (pairhmm_cuda.cu's one thread per pair,
    DP rows in global memory, FP32 with its 2^-40 rescale; accelerator.py's
    length buckets), batched, on a seeded random selection of whole regions
    of about `cells_target` cells; its NaN pairs are where that design falls
    back to FP64 (not run here)."""
    from vardepth_amd import engine
    rng = np.random.default_rng(12)
    order = rng.permutation(batch.nregions)
    rl, hl = batch.read_lengths(), batch.hap_lengths()
    take, total = [], 0
    for k in order:
        take.append(int(k))
        total += int(rl[batch.region_reads[k]:batch.region_reads[k + 1]].sum()) * \
            int(hl[batch.region_haps[k]:batch.region_haps[k + 1]].sum())
        if total >= cells_target:
            break
    sub = batch.subset(sorted(take))
    engine.forward_refdesign(sub.subset([0]), mode=mode)  # warm: module load, buffers
    out, ms, launches = engine.forward_refdesign(sub, mode=mode)
    cells = sub.cells()
    g = cells / ms / 1e6
    return {"gcups": round(g, 3), "regions": sub.nregions, "pairs": sub.npairs, "cells": cells,
            "gpu_ms": round(ms, 3), "launches": launches, "nan_frac": round(float(np.isnan(out).mean()), 4),
            "engine_speedup": round(engine_gcups / g, 2),
            "kind": "the reference's GPU design (one thread per pair, rows in global memory, FP32 + 2^-40 "
                    "rescale, 32-base length buckets), batched, inputs resident in HBM"}


def pair_lengths(batch):
    """Read and haplotype length of every pair, in the batch's output order
    (per region, row-major reads x haplotypes)."""
    rl, hl = batch.read_lengths(), batch.hap_lengths()
    ns, ms = [], []
    for k in range(batch.nregions):
        r = rl[batch.region_reads[k]:batch.region_reads[k + 1]]
        h = hl[batch.region_haps[k]:batch.region_haps[k + 1]]
        ns.append(np.repeat(r, len(h)))
        ms.append(np.tile(h, len(r)))
    return (np.concatenate(ns) if ns else np.zeros(0, np.int64)), (np.concatenate(ms) if ms else np.zeros(0, np.int64))


def accuracy_configs(mode: int):
    """The single 64x8 region and the 1024-region precision-stress workload:
    GCUPS, FP32 pairs rescued in FP64, and per-pair |log10 L - log10 L_FP64|
    (max, p99).  The FP64 reference of the stress set is the device FP64 path,
    itself checked against the CPU oracle on a 2-region sample; the single
    region is checked against the oracle directly."""
    import oracle
    from vardepth_amd import engine, workloads
    ln10 = math.log(10.0)
    out = {}
    b1 = workloads.single_region(seed=41001)
    engine.forward(b1, mode=mode)  # warm
    t0 = time.perf_counter()
    m1, st1 = engine.forward(b1, mode=mode, precision="mixed", stats=True)
    wall1 = time.perf_counter() - t0
    ref1, _ = oracle.regions_cpu(b1, mode=mode, threads=4)
    e1 = np.abs(m1 - ref1) / ln10
    out["single_region"] = {"workload": "single region 64 reads x 8 haplotypes (seed 41001, g=10^-4.5, x=0.1)",
                            "cells": b1.cells(), "device_ms": round(st1["device_ms"], 3),
                            "gcups_device": round(b1.cells() / st1["device_ms"] / 1e6, 2),
                            "wall_ms_with_transfers": round(wall1 * 1e3, 3),
                            "max_log10_err_vs_oracle": float(e1.max()), "flagged": int(st1["flagged"])}
    b3 = workloads.precision_stress(seed=41003, regions=1024)
    dev = engine.DeviceBatch(b3)
    dev.forward(mode=mode)
    st3 = dev.forward(mode=mode, precision="mixed", stats=True)
    m3 = dev.result()
    st64 = dev.forward(mode=mode, precision="float64", stats=True)
    f3 = dev.result()
    # the autotuner on this batch (launch choices and FP64 routing), then the
    # process's previous choices back for the legs that follow
    from vardepth_amd import autotune
    saved = autotune.get()
    try:
        tune3 = autotune.tune_forward(dev, mode=mode)
        st3t = min((dev.forward(mode=mode, precision="mixed", stats=True) for _ in range(3)),
                   key=lambda s_: s_["device_ms"])
        t3 = dev.result()
    finally:
        autotune.set_tuning(saved)
    okt = np.isfinite(f3)
    e3t = np.abs(t3[okt] - f3[okt]) / ln10
    ok = np.isfinite(f3)
    e3 = np.abs(m3[ok] - f3[ok]) / ln10
    # the reference's own rescue rule for its mixed GPU path (accelerator.py:139-154):
    # FP64 on the host when n > 512, m > 1024, the result is not finite or |ln L| > 80
    pn, pm = pair_lengths(b3)
    ref_rule = (pn > 512) | (pm > 1024) | ~np.isfinite(f3) | (np.abs(f3) > 80)
    sub = b3.subset([0, 1])
    ref3, _ = oracle.regions_cpu(sub, mode=mode, threads=4)
    n_sub = sub.npairs
    e_or = np.abs(f3[:n_sub] - ref3) / ln10
    out["precision_stress"] = {
        "workload": "precision stress (1024 regions, 16 x 250 bp reads Phred 2-15, 2-4 x 1-2 kb haplotypes with "
                    "an indel per 50 bp, seed 41003)",
        "pairs": b3.npairs, "cells": b3.cells(), "fwd_ms": round(st3["device_ms"], 3),
        "gcups": round(b3.cells() / st3["device_ms"] / 1e6, 2), "fp64_all_ms": round(st64["device_ms"], 3),
        "flagged": int(st3["flagged"]), "flagged_frac": round(st3["flagged"] / b3.npairs, 4),
        "reference_rule_frac": round(float(ref_rule.mean()), 4),
        "tuned_ms": round(st3t["device_ms"], 3), "tuned_gcups": round(b3.cells() / st3t["device_ms"] / 1e6, 2),
        "tuned_route64_classes": tune3.get("route64", {}).get("classes", []),
        "tuned_max_log10_err_vs_fp64": float(e3t.max()),
        "reference_rule_abs_ln_gt_80_frac": round(float((np.abs(f3) > 80).mean()), 4),
        "max_log10_err_vs_fp64": float(e3.max()), "p99_log10_err_vs_fp64": float(np.quantile(e3, 0.99)),
        "fp64_vs_cpu_oracle_max_log10": float(e_or.max()), "oracle_sample_pairs": int(n_sub)}
    return out


def grad_bench(regions: int, steps: int, mode: int, tune: bool = True, train_steps: int = 200):
    """Forward and forward+adjoint GCUPS on the HMM-sampled workload (all pairs,
    unit weights: gradient of sum ln L wrt gap, ext and the 64-bin error table).

    `fwd_bwd_gcups` / `ratio` count a forward pass plus a gradient call, as an
    autograd step runs them.  One gradient call also returns ln L per pair
    (its first sweep is the forward, and the FP64 adjoint rescues the pairs
    it flags), so `fused_gcups` / `fused_ratio` count that call alone.
    `training` is the differentiation workload's own run: train_steps (200) Adam steps on the
    true-haplotype loss from a perturbed start, timed, with the recovered
    parameters (None when train_steps is 0)."""
    import torch
    from vardepth_amd import engine, workloads
    b = workloads.hmm_sampled_regions(seed=41004, regions=regions)
    cells = b.cells()
    dev = engine.DeviceBatch(b)
    tune_f = tune_g = None
    if tune:
        from vardepth_amd import autotune
        tune_f = autotune.tune_forward(dev, mode=mode)
        tune_g = autotune.tune_grad(dev, mode=mode)
    dev.forward(mode=mode)
    dev.grad(mode=mode)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.forward(mode=mode)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) / steps * 1e3
    # the same forward without haplotype prefix sharing (the adjoint sweeps
    # every pair in full): the like-for-like denominator
    from vardepth_amd import _lib
    prev = _lib.lib().vd_prefix_sharing(0)
    try:
        dev.forward(mode=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            dev.forward(mode=mode)
        torch.cuda.synchronize()
        fwd_full_ms = (time.perf_counter() - t0) / steps * 1e3
    finally:
        _lib.lib().vd_prefix_sharing(prev)
    sts = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sts.append(dev.grad(mode=mode, stats=True)[1])
    torch.cuda.synchronize()
    grad_ms = (time.perf_counter() - t0) / steps * 1e3
    k32 = float(np.mean([s_["fwd32_ms"] for s_ in sts]))
    # the same call with weight 1 on each read's true haplotype and 0 elsewhere
    # (a loss on the true pairs of the full matrix): units whose four weights
    # are zero write their ln L and skip the backward sweeps
    oo = b.out_off()
    w = np.zeros(b.npairs)
    for k in range(b.nregions):
        r0, r1, h0 = b.region_reads[k], b.region_reads[k + 1], b.region_haps[k]
        H = b.region_haps[k + 1] - h0
        w[oo[k] + np.arange(r1 - r0) * H + (b.read_true_hap[r0:r1] - h0)] = 1.0
    wdev = torch.from_numpy(w).to("cuda")
    dev.grad(mode=mode, weights=wdev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.grad(mode=mode, weights=wdev)
    torch.cuda.synchronize()
    truth_w_ms = (time.perf_counter() - t0) / steps * 1e3
    # the training run: the loss -mean ln L over each read's true
    # haplotype (the true-pair batch), train_steps Adam steps from a perturbed
    # start, each step one forward and one adjoint call through torch.autograd
    training = None
    if train_steps > 0:
        from vardepth_amd import fit as fitmod
        true_gap, true_ext = 0.01, 0.3
        err_true = np.power(10.0, -np.minimum(np.arange(64), 60) / 10.0)
        tp = workloads.true_pairs(b)
        devt = engine.DeviceBatch(tp)
        tune_t = None
        if tune:  # the true-pair batch's own launch choices (its regions hold few pairs)
            from vardepth_amd import autotune
            tune_t = {"forward": autotune.tune_forward(devt, mode=mode), "adjoint": autotune.tune_grad(devt, mode=mode)}
        gap0, ext0, err0 = fitmod.perturbed_start(41004, true_gap, true_ext, err_true)
        fitmod.fit(devt, gap0, ext0, err0, steps=2, report=2)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, hist = fitmod.fit(devt, gap0, ext0, err0, steps=train_steps, lr=0.05, report=train_steps)
        torch.cuda.synchronize()
        fit_s = time.perf_counter() - t0
        used = np.arange(10, 41)
        lerr = lambda e: float(np.mean(np.abs(np.log10(np.asarray(e)[used]) - np.log10(err_true[used]))))
        training = {"workload": "true-haplotype pairs of the same regions (loss -mean ln L)", "pairs": tp.npairs,
                    "cells": tp.cells(), "adam_steps": train_steps, "seconds": round(fit_s, 3),
                    "ms_per_step": round(fit_s / train_steps * 1e3, 3),
                    "fwd_bwd_gcups": round(tp.cells() * train_steps / fit_s / 1e9, 2),
                    "gap": [round(hist[0]["gap"], 5), round(hist[-1]["gap"], 5), true_gap],
                    "ext": [round(hist[0]["ext"], 4), round(hist[-1]["ext"], 4), true_ext],
                    "mean_abs_log10_err_q10_40": [round(lerr(hist[0]["err"]), 4), round(lerr(hist[-1]["err"]), 4)],
                    "loss": [round(hist[0]["loss"], 4), round(hist[-1]["loss"], 4)], "autotune": tune_t}
    fwd_g = cells / fwd_ms / 1e6
    fb_g = cells / (fwd_ms + grad_ms) / 1e6
    fused_g = cells / grad_ms / 1e6
    return {"workload": "hmm-sampled (%d regions x 100 reads x 20 haplotypes, 150 bp, seed 41004)" % regions,
            "pairs": b.npairs, "cells": cells, "fwd_ms": round(fwd_ms, 3), "adjoint_ms": round(grad_ms, 3),
            "adjoint_kernel_ms": round(k32, 3), "fwd_gcups": round(fwd_g, 3), "fwd_bwd_gcups": round(fb_g, 3),
            "ratio": round(fb_g / fwd_g, 4), "fused_gcups": round(fused_g, 3),
            "fused_ratio": round(fused_g / fwd_g, 4),
            "fwd_unshared_ms": round(fwd_full_ms, 3), "fused_ratio_vs_unshared_fwd": round(fwd_full_ms / grad_ms, 4),
            "truth_weights_adjoint_ms": round(truth_w_ms, 3),
            "flagged": int(sts[-1]["flagged"]),
            "autotune": {"forward": tune_f, "adjoint": tune_g}, "training": training}


def scalar_bench(pairs: int = 2000, calls: int = 400):
    """Per-call latency of the scalar drop-ins on the short anchored windows the
    reference scores one read at a time (indel_refine.py:160-163,
    partial_anchor.py:89-97: a cropped read against an allele with 3 bp of
    flank): anchored_pairhmm and wide read_haplotype of this library (one
    device call per pair) against the reference's scalar C++ built into
    oracle/_ref (one core), and the batched drop-in (cpu_batch.pairs(mode=1,
    canonical=True): all pairs in one device call) for the same pairs."""
    import ctypes
    import oracle
    from vardepth_amd import batch, native
    rng = np.random.default_rng(13)
    jobs = []
    for _ in range(pairs):
        n = int(rng.integers(20, 81))
        s = bytes(rng.choice(list(b"ACGT"), n).astype(np.uint8))
        cut = int(rng.integers(3, n - 3))
        h = s[:cut] + bytes(rng.choice(list(b"ACGT"), int(rng.integers(0, 6))).astype(np.uint8)) + s[cut + int(rng.integers(0, 3)):]
        q = bytes(rng.integers(10, 41, n).astype(np.uint8))
        jobs.append((s, q, h))
    qbufs = [(ctypes.c_ubyte * len(q)).from_buffer_copy(q) for _, q, _ in jobs]
    out = {"workload": "%d cropped reads of 20-80 bp against their allele windows (0-5 bp indel), Phred 10-40"
                       % pairs}
    fns = {"anchored_pairhmm": (native.kernel().anchored_pairhmm, "anchored"),
           "read_haplotype": (native.wide_library().read_haplotype, "wide")}
    ref = oracle.ref()
    for name, (fn, refname) in fns.items():
        fn(jobs[0][0], qbufs[0], len(jobs[0][0]), jobs[0][2], len(jobs[0][2]), .001, .1)   # warm
        t0 = time.perf_counter()
        for i in range(calls):
            s, _, h = jobs[i % pairs]
            fn(s, qbufs[i % pairs], len(s), h, len(h), .001, .1)
        gpu_us = (time.perf_counter() - t0) / calls * 1e6
        rec = {"gpu_us_per_call": round(gpu_us, 2)}
        if ref is not None:
            rf = {"anchored": ref.anchored.anchored_pairhmm, "wide": ref.wide.read_haplotype}[refname]
            t0 = time.perf_counter()
            for i in range(pairs):
                s, _, h = jobs[i]
                rf(s, qbufs[i], len(s), h, len(h), .001, .1)
            rec["cpu_reference_us_per_call"] = round((time.perf_counter() - t0) / pairs * 1e6, 3)
        out[name] = rec
    strjobs = [(s.decode(), q, h.decode()) for s, q, h in jobs]

    def batched(k):
        batch._CACHE.clear()
        t0 = time.perf_counter()
        batch.pairs(strjobs[:k], mode=1, canonical=True)
        return time.perf_counter() - t0

    batched(16)   # warm
    t_small = min(batched(16) for _ in range(5))
    t_full = min(batched(pairs) for _ in range(3))
    per_pair = (t_full - t_small) / (pairs - 16)
    fixed = max(t_small - 16 * per_pair, 0.0)
    out["batched_pairs_mode1"] = {"pairs": pairs, "call_ms": round(t_full * 1e3, 3),
                                  "us_per_pair": round(t_full / pairs * 1e6, 3),
                                  "call_16_pairs_ms": round(t_small * 1e3, 3)}
    cpu = out["anchored_pairhmm"].get("cpu_reference_us_per_call")
    if cpu and cpu * 1e-6 > per_pair:
        # a batched call of N pairs (fixed + N per_pair) beats N scalar calls on one CPU core from N on
        out["batched_break_even_pairs"] = int(math.ceil(fixed / (cpu * 1e-6 - per_pair)))
    return out


def genotype_bench(mb: float, depths, mode: int, steps: int, cpu_seconds: float, threads: int, tune: bool = True):
    """Forward + diploid genotype likelihood on the chromosome-scale sample
    (launch choices tuned on the sample itself unless tune is False)."""
    import torch
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from vardepth_amd import engine, workloads
    ref = oracle.ref()
    res = {}
    for depth in depths:
        t0 = time.perf_counter()
        # the region call's CPU path (pair_scores_many -> cpu_batch -> trie_many)
        # computes at g = .001, x = .1, so the sample is scored at those values
        cs = workloads.chromosome_sample(mb=mb, depth=depth, gap=REFERENCE_GAP_EXT[0], ext=REFERENCE_GAP_EXT[1],
                                         progress=lambda m: print(m, file=sys.stderr, flush=True))
        gen_s = time.perf_counter() - t0
        b = cs.batch
        cells = b.cells()
        nh = np.diff(b.region_haps)
        gl_off = engine.gl_offsets(nh)
        dev = engine.DeviceBatch(b)
        mq = torch.from_numpy(cs.mq).cuda()
        tuning = None
        if tune:
            from vardepth_amd import autotune
            tuning = autotune.tune_forward(dev, mode=mode)
        dev.forward(mode=mode)
        dev.genotype(mq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            dev.forward(mode=mode)
            gl = dev.genotype(mq)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / steps
        fst = dev.forward(mode=mode, stats=True)
        gl, _, cert, gst = dev.genotype(mq, stats=True, values=True)
        g = gl.cpu().numpy()
        ll = dev.result()
        # k_mixture's algorithmic work: every kept row x pair x mixture term, 4 FP64
        # flops each (two products and a sum or max forming the mixture value, the
        # running product); the clean-only / no-mixture priors are not used here
        kept = (~(cert.cpu().numpy() < engine.GENO_MIN_CERTAINTY)).astype(np.int64)
        csum = np.concatenate([[0], np.cumsum(kept)])
        kept_k = csum[b.region_reads[1:]] - csum[b.region_reads[:-1]]
        terms = int(np.sum(kept_k * (nh * (nh + 1) // 2) * (7 + 4 * nh)))
        mix_tflops = terms * 4 / (gst["mixture_ms"] * 1e-3) / 1e12 if gst["mixture_ms"] > 0 else 0.0
        calls = np.array([int(np.argmax(g[gl_off[k]:gl_off[k + 1]])) for k in range(b.nregions)])
        var = nh >= 2
        # CPU chain on a time-bounded random sample of regions
        rng = np.random.default_rng(5)
        order = rng.permutation(b.nregions)
        done, cpu_cells, cpu_t, maxdiff, stage_diff, agree, nvar = 0, 0, 0.0, 0.0, 0.0, 0, 0
        while done < b.nregions and cpu_t < cpu_seconds:
            ks = np.sort(order[done:done + 128])
            done += len(ks)
            sub = b.subset(ks)
            snh = np.diff(sub.region_haps).astype(np.int32)
            soff = engine.gl_offsets(snh)
            smq = np.concatenate([cs.mq[b.region_reads[k]:b.region_reads[k + 1]] for k in ks])
            t1 = time.perf_counter()
            if ref is not None:
                with ThreadPoolExecutor(max_workers=threads) as ex:
                    sll = np.concatenate(list(ex.map(lambda k: reference_region(ref, b, int(k), mode), ks)))
            else:
                sll, _ = oracle.regions_cpu(sub, mode=mode, threads=threads)
            sgl, _ = oracle.genotype_regions(sub.region_reads, snh, sub.out_off()[:-1], sll, smq, soff,
                                             threads=threads)
            cpu_t += time.perf_counter() - t1
            cpu_cells += sub.cells()
            # the device's genotype stage vs the CPU chain on the device's own likelihoods
            oo = b.out_off()
            dll = np.concatenate([ll[oo[k]:oo[k + 1]] for k in ks])
            dgl, _ = oracle.genotype_regions(sub.region_reads, snh, sub.out_off()[:-1], dll, smq, soff,
                                             threads=threads)
            for i, k in enumerate(ks):
                mine = g[gl_off[k]:gl_off[k + 1]]
                maxdiff = max(maxdiff, float(np.max(np.abs(mine - sgl[soff[i]:soff[i + 1]]))))
                stage_diff = max(stage_diff, float(np.max(np.abs(mine - dgl[soff[i]:soff[i + 1]]) /
                                                          np.maximum(1, np.abs(dgl[soff[i]:soff[i + 1]])))))
                if snh[i] >= 2:
                    nvar += 1
                    agree += int(np.argmax(sgl[soff[i]:soff[i + 1]]) == calls[k])
        cpu_gcups = cpu_cells / cpu_t / 1e9 if cpu_t > 0 else None
        res["%gx" % depth] = {
            "workload": "chromosome sample %.3g Mb at %gx (seed 41005, g=%.3g, x=%.3g): %d regions, %d haplotypes "
                        "(max %d), %d reads in regions" % (mb, depth, b.gap, b.ext, b.nregions, b.nhaps,
                                                           int(nh.max()), b.nreads),
            "regions": b.nregions, "pairs": b.npairs, "genotype_outputs": int(gl_off[-1]), "cells": cells,
            "generate_s": round(gen_s, 1),
            "end_to_end_ms": round(wall * 1e3, 3), "end_to_end_gcups": round(cells / wall / 1e9, 2),
            "forward_ms": round(fst["device_ms"], 3), "fp32_cells_swept": int(fst["fp32_cells_swept"]),
            "autotune": tuning, "molecule_ms": round(gst["molecule_ms"], 3),
            "mixture_ms": round(gst["mixture_ms"], 3), "mixture_units": gst["units"],
            "kept_rows": gst["kept_rows"], "invalid_regions": gst["invalid"],
            "mixture_roofline": {"bound": "valu-fp64", "terms": terms, "flops_per_term": 4,
                                 "achieved": round(mix_tflops, 3), "peak": 78.6, "unit": "TFLOP/s",
                                 "frac": round(mix_tflops / 78.6, 4)},
            "truth_concordance": round(float(np.mean(calls[var] == cs.true_pair[var])), 4) if var.any() else None,
            "cpu": {"gcups": round(cpu_gcups, 3) if cpu_gcups else None, "threads": threads,
                    "cores_available": available_cpus(),
                    "kind": "reference" if ref is not None and (b.gap, b.ext) == REFERENCE_GAP_EXT else "port",
                    "sample_regions": int(done),
                    "sample_cells": int(cpu_cells), "seconds": round(cpu_t, 2),
                    "projected_wall_s": round(cells / (cpu_gcups * 1e9), 1) if cpu_gcups else None,
                    "what": ("trie_many + pairhmm_pairs compiled from the reference sources, one region per "
                             "thread" if ref is not None else "oracle regions_cpu (C restatement of trie_many + "
                             "pairhmm_pairs)") + ", then the oracle genotype chain (C restatement of "
                             "molecule_rows / mixture_two / mixture_general)"},
            "speedup_vs_cpu": round(cells / wall / 1e9 / cpu_gcups, 1) if cpu_gcups else None,
            "gl_max_abs_diff_vs_cpu_chain": maxdiff,
            "genotype_stage_max_rel_diff": stage_diff,
            "call_agreement_vs_cpu": round(agree / nvar, 4) if nvar else None,
        }
        print("genotype %gx: %s" % (depth, json.dumps(res["%gx" % depth])), file=sys.stderr, flush=True)
        del dev, gl, mq
        torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--regions", type=int, default=4096)
    ap.add_argument("--workload", default="uneven", choices=["uneven", "single", "stress"])
    ap.add_argument("--mode", type=int, default=0, choices=[0, 1, 2])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "float64"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch choices)")
    ap.add_argument("--tune-save", help="after tuning, keep the choice in this JSON file (autotune.save)")
    ap.add_argument("--tune-load", help="instead of tuning, apply the choice kept in this JSON file")
    ap.add_argument("--no-forward", action="store_true",
                    help="skip the forward workload (profiling the fwd+bwd workload alone); prints no JSON line")
    ap.add_argument("--plain-pairs", type=int, default=20000, help="pairs of the plain baseline sample (0 = skip)")
    ap.add_argument("--ref-cells", type=float, default=5e10,
                    help="cells of the reference-GPU-design baseline sample (0 = skip)")
    ap.add_argument("--grad-regions", type=int, default=1024, help="regions of the fwd+bwd workload (0 = skip)")
    ap.add_argument("--grad-steps", type=int, default=5)
    ap.add_argument("--train-steps", type=int, default=200,
                    help="Adam steps of the gradient leg's training run (0 = skip)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the single-region and precision-stress reports")
    ap.add_argument("--no-scalar", action="store_true", help="skip the scalar drop-in latency leg")
    ap.add_argument("--geno-mb", type=float, default=60.0,
                    help="chromosome-sample size for the genotype leg, Mb (0 = skip; default: the full 60 Mb sample)")
    ap.add_argument("--geno-depths", default="30,2", help="depth profiles of the genotype leg")
    ap.add_argument("--geno-cpu-seconds", type=float, default=10.0)
    ap.add_argument("--only-genotype", action="store_true", help="run only the genotype leg")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one 256 MiB device-to-device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the regions shard across GPUs with no exchange, "
              "so only --gpus 1 is run here (asked for %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from vardepth_amd import engine

    if args.copy_calib and (args.only_genotype or args.no_forward):
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    if args.only_genotype:
        print(json.dumps({"genotype": genotype_bench(args.geno_mb, [float(d) for d in args.geno_depths.split(",")],
                                                     args.mode, max(1, min(args.steps, 3)), args.geno_cpu_seconds,
                                                     min(available_cpus(), CPU_THREAD_CAP), not args.no_tune)}), flush=True)
        return

    if args.no_forward:
        fb = grad_bench(args.grad_regions, args.grad_steps, args.mode, not args.no_tune, args.train_steps)
        print(json.dumps({"fwd_bwd": fb}), flush=True)
        return

    t0 = time.perf_counter()
    batch, wname = workload(args.workload, args.regions)
    gen_s = time.perf_counter() - t0
    cells = batch.cells()
    cells32 = fp32_cells(batch, args.mode)
    print("workload %s: %d regions %d pairs %.4g cells (generated in %.1fs)" % (
        wname, batch.nregions, batch.npairs, cells, gen_s), file=sys.stderr, flush=True)

    dev = engine.DeviceBatch(batch)
    tune = None
    tune_key = "%s-%d-mode%d-%s" % (args.workload, args.regions, args.mode, args.precision)
    if args.tune_load:
        from vardepth_amd import autotune
        if not autotune.load(args.tune_load, tune_key):
            raise SystemExit("no tuning for %s in %s" % (tune_key, args.tune_load))
        tune = {"kind": "forward", "loaded": args.tune_load}
    elif not args.no_tune:
        from vardepth_amd import autotune
        t0 = time.perf_counter()
        tune = autotune.tune_forward(dev, mode=args.mode, precision=args.precision)
        tune["seconds"] = round(time.perf_counter() - t0, 2)
        print("autotune forward: %s" % json.dumps(tune), file=sys.stderr, flush=True)
        if args.tune_save:
            autotune.save(args.tune_save, tune_key)
    for _ in range(args.warmup):
        dev.forward(mode=args.mode, precision=args.precision, stats=True)
    torch.cuda.synchronize()
    steps = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        steps.append(dev.forward(mode=args.mode, precision=args.precision, stats=True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    gcups = cells * args.steps / wall / 1e9
    headline_out = dev.result()   # the last timed step's likelihoods, checked against the CPU reference below
    fwd32 = float(np.mean([s["fwd32_ms"] for s in steps]))
    last = steps[-1]
    # the FP32 kernels' hardware rate counts the cells they computed (fewer than
    # the pairs' n*m when haplotype prefixes are shared); the rate over the
    # pairs' n*m is the effective one that GCUPS reports
    swept = int(last.get("fp32_cells_swept", 0)) or cells32
    achieved = swept * FLOPS_PER_CELL / (fwd32 * 1e-3) / 1e12 if fwd32 > 0 else 0.0
    effective = cells32 * FLOPS_PER_CELL / (fwd32 * 1e-3) / 1e12 if fwd32 > 0 else 0.0
    print("step %.2f ms: plan %.2f fwd32 %.2f fwd64 %.2f ms; flagged %d, fp64 pairs %d" % (
        ms_step, last["plan_ms"], last["fwd32_ms"], last["fwd64_ms"], last["flagged"], last["fp64_pairs"]),
        file=sys.stderr, flush=True)

    graph = None
    if not args.no_graph:
        g = dev.graph(mode=args.mode, precision=args.precision)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        gms = (time.perf_counter() - t0) / args.steps * 1e3
        graph = {"ms_per_step": round(gms, 3), "gcups": round(cells / gms / 1e6, 3), "replays": args.steps}
        print("graph replay %.2f ms/step" % gms, file=sys.stderr, flush=True)
        del g

    plain = None
    if args.plain_pairs > 0:
        plain = plain_baseline(batch, args.mode, args.plain_pairs)
        plain["speedup"] = round(gcups / plain["gcups"], 1)
        print("plain one-pair-per-launch: %.3f GCUPS on %d pairs" % (plain["gcups"], plain["pairs"]),
              file=sys.stderr, flush=True)

    ref_design = None
    if args.ref_cells > 0:
        ref_design = refdesign_baseline(batch, args.mode, args.ref_cells, gcups)
        print("reference GPU design: %.2f GCUPS on %d pairs" % (ref_design["gcups"], ref_design["pairs"]),
              file=sys.stderr, flush=True)

    accuracy = None
    if not args.no_accuracy:
        accuracy = accuracy_configs(args.mode)
        print("accuracy: %s" % json.dumps(accuracy), file=sys.stderr, flush=True)

    fwd_bwd = None
    if args.grad_regions > 0:
        fwd_bwd = grad_bench(args.grad_regions, args.grad_steps, args.mode, not args.no_tune, args.train_steps)
        print("fwd+bwd: %s" % json.dumps(fwd_bwd), file=sys.stderr, flush=True)

    scalar = None
    if not args.no_scalar:
        scalar = scalar_bench()
        print("scalar drop-ins: %s" % json.dumps(scalar), file=sys.stderr, flush=True)

    genotype = None
    if args.geno_mb > 0:
        genotype = genotype_bench(args.geno_mb, [float(d) for d in args.geno_depths.split(",")], args.mode,
                                  max(1, min(args.steps, 3)), args.geno_cpu_seconds, min(available_cpus(), CPU_THREAD_CAP),
                                  not args.no_tune)

    if args.copy_calib:
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu:
        threads = min(available_cpus(), CPU_THREAD_CAP)
        cpu = cpu_baseline(batch, args.mode, args.cpu_seconds, threads, device_out=headline_out)

    line = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": wname, "regions": batch.nregions, "pairs": batch.npairs, "cells": cells,
                   "mode": args.mode, "precision": "fp32 bulk + fp64 rescue" if args.precision == "mixed" else "fp64",
                   "gap": batch.gap, "ext": batch.ext,
                   "flagged_pairs": last["flagged"], "fp64_pairs": last["fp64_pairs"],
                   "stage_ms": {"plan": round(last["plan_ms"], 3), "fwd32": round(last["fwd32_ms"], 3),
                                "fwd64": round(last["fwd64_ms"], 3)},
                   "autotune": tune},
        "roofline": {"bound": "valu", "kernel": "k_fwd32", "achieved": round(achieved, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic(args), "traffic_unit": "bytes/launch (rocprofv3 PMC)",
                     "flops_per_cell": FLOPS_PER_CELL, "cells_per_launch": swept,
                     "pair_cells_per_launch": cells32, "effective": round(effective, 3),
                     "effective_frac": round(effective / FP32_PEAK_TFLOPS, 4),
                     "kernel_ms": round(fwd32, 3)},
        "cpu_baseline": cpu,
        "graph": graph,
        "plain": plain,
        "ref_design": ref_design,
        "fwd_bwd": fwd_bwd,
        "accuracy": accuracy,
        "scalar_dropins": scalar,
        "genotype": genotype,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
