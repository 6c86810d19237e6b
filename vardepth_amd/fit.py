"""Recover PairHMM parameters by gradient descent on the true-haplotype likelihood.

    python -m vardepth_amd.fit [--regions 1024] [--steps 200]

Reads are sampled from the PairHMM itself (vardepth_amd.workloads.
hmm_sampled_regions: gap 0.01, continuation 0.3, error 10^(-q/10) for the
sampled qualities 10..40); the loss is -mean_r ln L(read r, its true
haplotype); the parameters (gap open, gap continuation, 64-bin error table,
FP32) start from a perturbed point and follow Adam through the HIP forward and
adjoint kernels (vardepth_amd.autograd).  Prints one JSON line per report
step and a summary line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def perturbed_start(seed: int, gap: float, ext: float, err):
    rng = np.random.default_rng(seed)
    return gap * 3.0, ext * 0.5, np.clip(np.asarray(err) * np.exp(rng.normal(0, 0.7, 64)), 1e-7, 0.7)


def fit(dev_batch, gap0, ext0, err0, steps=200, lr=0.05, mode=0, report=20, log=None):
    import torch
    from .autograd import HmmParameters
    params = HmmParameters(gap0, ext0, err0)
    opt = torch.optim.Adam(params.parameters(), lr=lr)
    history = []
    n = dev_batch.npairs
    for step in range(steps + 1):
        opt.zero_grad()
        ll = params(dev_batch, mode, fused=step < steps)  # the loss -mean ln L: one adjoint call per step
        loss = -ll.sum() / n
        if step % report == 0 or step == steps:
            g, x, e = params.values()
            rec = {"step": step, "loss": loss.item(), "gap": g.item(), "ext": x.item(),
                   "err": e.detach().cpu().numpy().tolist()}
            history.append(rec)
            if log:
                log(rec)
        if step == steps:
            break
        loss.backward()
        opt.step()
    return params, history


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--regions", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--lr", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=41004)
    args = ap.parse_args()
    from . import engine, workloads
    true_gap, true_ext = 0.01, 0.3
    err_true = np.power(10.0, -np.minimum(np.arange(64), 60) / 10.0)
    b = workloads.hmm_sampled_regions(seed=args.seed, regions=args.regions, sample_gap=true_gap,
                                      sample_ext=true_ext)
    tp = workloads.true_pairs(b)
    dev = engine.DeviceBatch(tp)
    gap0, ext0, err0 = perturbed_start(args.seed, true_gap, true_ext, err_true)
    t0 = time.perf_counter()
    params, hist = fit(dev, gap0, ext0, err0, steps=args.steps, lr=args.lr,
                       log=lambda r: print(json.dumps({k: v for k, v in r.items() if k != "err"}), flush=True))
    dt = time.perf_counter() - t0
    used = np.arange(10, 41)
    last, first = hist[-1], hist[0]
    lerr = lambda e: float(np.mean(np.abs(np.log10(np.asarray(e)[used]) - np.log10(err_true[used]))))
    print(json.dumps({"pairs": tp.npairs, "cells": tp.cells(), "steps": args.steps, "seconds": round(dt, 2),
                      "gap": [first["gap"], last["gap"], true_gap], "ext": [first["ext"], last["ext"], true_ext],
                      "mean_abs_log10_err_q10_40": [lerr(first["err"]), lerr(last["err"])],
                      "loss": [first["loss"], last["loss"]]}), flush=True)


if __name__ == "__main__":
    main()
