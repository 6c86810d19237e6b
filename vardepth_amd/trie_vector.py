"""Read-grouped region batches on the MI355X: the host interface of ``trie_vector.py``.

:func:`score` keeps the reference's grouping and routing (trie_vector.py:16-42):
jobs are grouped by read ``(s, q)``, reads that share an identical sorted
haplotype set form one batch (reads sorted by length, haplotypes sorted),
every value at or below 1e-200 (the reference's trie underflow rule,
trie_vector.cpp:277) becomes NaN and NaN values are recomputed by ``dense``.

Unlike the reference, which calls ``trie_many`` once per batch, all batches
of a call go to the device together as the regions of one batched launch.
Precision follows ``VARDEPTH_CUDA_PRECISION`` (default ``mixed``: FP32 with
FP64 rescue, <= 1e-4 log10 per pair; ``float64``: every pair in FP64).
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import engine
from .native import trie_library as library  # noqa: F401  (reference name)
from .workloads import from_regions

LN_TRIE_FLOOR = math.log(1e-200)


def precision() -> str:
    value = os.environ.get("VARDEPTH_CUDA_PRECISION", "mixed")
    if value not in ("float64", "mixed"):
        raise ValueError("CUDA precision must be float64 or mixed")
    return value


def score(jobs, mode, dense, lanes=0):
    """(P,) ln likelihoods; ``dense(jobs, mode, lanes)`` recomputes the NaN ones."""
    if not jobs:
        return np.empty(0)
    grouped = {}
    qualities = {}
    for index, (s, q, h) in enumerate(jobs):
        if len(s) != len(q):
            raise ValueError('Read and quality lengths differ')
        if id(q) not in qualities:
            qualities[id(q)] = q if isinstance(q, bytes) else np.ascontiguousarray(q, np.uint8).tobytes()
        grouped.setdefault((s, qualities[id(q)]), {}).setdefault(h, []).append(index)
    batches = {}
    for key, haps in grouped.items():
        batches.setdefault(tuple(sorted(haps)), []).append(key)
    regions, layout = [], []
    for haps, reads in batches.items():
        reads.sort(key=lambda x: len(x[0]))
        regions.append(([s for s, q in reads], [q for s, q in reads], list(haps)))
        layout.append((reads, haps))
    batch = from_regions(regions, 0.001, 0.1)
    out = engine.forward(batch, mode=mode, precision=precision())
    out = np.where(np.isinf(out) | (out > LN_TRIE_FLOOR), out, np.nan)
    oo = batch.out_off()
    result = np.empty(len(jobs))
    for k, (reads, haps) in enumerate(layout):
        block = out[oo[k]:oo[k + 1]].reshape(len(reads), len(haps))
        for ri, key in enumerate(reads):
            for hi, h in enumerate(haps):
                for j in grouped[key][h]:
                    result[j] = block[ri, hi]
    fallback = np.flatnonzero(np.isnan(result))
    if len(fallback):
        result[fallback] = dense([jobs[j] for j in fallback], mode, 0)
    return result
