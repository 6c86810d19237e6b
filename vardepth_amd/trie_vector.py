"""Read-grouped region batches on the MI355X: the interface of ``trie_vector.py``.

:func:`score` has the reference's arguments, grouping and routing
(trie_vector.py:16-42): reads are identified by ``(sequence, quality)``;
reads whose sets of haplotypes are identical form one region, with the reads
in order of length and the haplotypes in sorted order; a value at or below
1e-200 -- where the reference's trie reports NaN (trie_vector.cpp:277) --
becomes NaN, and NaN values are recomputed by ``dense(jobs, mode, 0)``.

The reference calls ``trie_many`` once per region.  Here the job list is
interned (:class:`~vardepth_amd.jobs.JobTable`), every region of the call
goes to the device in one batched launch, and the values are gathered back to
job order by one index computation: job j sits in row ``row[read(j)]`` and
column ``col[hap(j)]`` of its region's block.  Precision follows
``VARDEPTH_CUDA_PRECISION`` (default ``mixed``: FP32 with FP64 rescue,
<= 1e-4 log10 per pair; ``float64``: every pair in FP64).
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import engine
from .jobs import JobTable
from .native import trie_library as library  # noqa: F401  (reference name)

LN_TRIE_FLOOR = math.log(1e-200)


def precision() -> str:
    value = os.environ.get("VARDEPTH_CUDA_PRECISION", "mixed")
    if value not in ("float64", "mixed"):
        raise ValueError("CUDA precision must be float64 or mixed")
    return value


def _read_groups(table: JobTable):
    """Regions of identical haplotype sets.

    Returns (read id per job, the regions as (read ids, haplotype ids sorted by
    bytes)), read ids numbered in first-occurrence order."""
    _, first, read_of = np.unique(table.read_key(), return_index=True, return_inverse=True)
    read_of = read_of.reshape(-1)
    renum = np.empty(len(first), np.int64)
    renum[np.argsort(first, kind="stable")] = np.arange(len(first))
    read_of = renum[read_of]
    # haplotype rank in byte order: sets become sorted tuples of ranks
    by_bytes = sorted(range(len(table.haps)), key=table.haps.__getitem__)
    rank = np.empty(len(by_bytes), np.int64)
    rank[by_bytes] = np.arange(len(by_bytes))
    pair = np.unique(read_of * max(1, len(by_bytes)) + rank[table.hap])
    owner, hrank = np.divmod(pair, max(1, len(by_bytes)))
    bounds = np.flatnonzero(np.diff(owner)) + 1
    sets = {}
    members = []
    for r, ranks in zip(owner[np.r_[0, bounds]] if len(owner) else [], np.split(hrank, bounds)):
        key = ranks.tobytes()
        k = sets.get(key)
        if k is None:
            k = sets[key] = len(members)
            members.append(([], ranks))
        members[k][0].append(int(r))
    by_bytes = np.asarray(by_bytes, np.int64)
    return read_of, [(reads, by_bytes[ranks]) for reads, ranks in members]


def score(jobs, mode, dense, lanes=0):
    """(P,) ln likelihoods; ``dense(jobs, mode, lanes)`` recomputes the NaN ones."""
    if not jobs:
        return np.empty(0)
    table = JobTable(jobs)
    read_of, regions = _read_groups(table)
    first_job = np.empty(read_of.max() + 1, np.int64)
    first_job[read_of[::-1]] = np.arange(len(read_of))[::-1]
    rlen = np.fromiter((len(table.seqs[s]) for s in table.seq[first_job]), np.int64, len(first_job))
    read_jobs, hap_ids = [], []
    row = np.empty(len(first_job), np.int64)
    region_of = np.empty(len(first_job), np.int64)
    col = {}
    for k, (reads, haps) in enumerate(regions):
        reads = sorted(reads, key=rlen.__getitem__)             # stable: equal lengths keep input order
        row[reads] = np.arange(len(reads))
        region_of[reads] = k
        read_jobs.append(first_job[reads])
        hap_ids.append(haps)
        for c, h in enumerate(haps):
            col[k, int(h)] = c
    b = table.region_batch(read_jobs, hap_ids)
    out = engine.forward(b, mode=mode, precision=precision())
    out = np.where(np.isinf(out) | (out > LN_TRIE_FLOOR), out, np.nan)
    oo = b.out_off()
    width = np.diff(b.region_haps)
    k_of = region_of[read_of]
    cols = np.fromiter((col[int(k), int(h)] for k, h in zip(k_of, table.hap)), np.int64, table.count)
    result = out[oo[k_of] + row[read_of] * width[k_of] + cols]
    fallback = np.flatnonzero(np.isnan(result))
    if len(fallback):
        result[fallback] = dense([jobs[j] for j in fallback], mode, 0)
    return result
