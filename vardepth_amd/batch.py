"""Cached pair batches on the MI355X: the interface of ``cpu_batch.py``.

:func:`pairs`, :func:`many` and :func:`_uncached_pairs` take the reference's
arguments and raise its errors (cpu_batch.py:20-66).  Behaviour kept:

* a likelihood is a pure function of (sequence, quality, haplotype, mode), so
  values are cached under ``(mode, lanes, s, q, h, canonical)`` -- ``q`` as
  bytes -- in a least-recently-used table of ``_CACHE_LIMIT`` = 131,072
  entries, and a job repeated within one call is computed once;
* ``canonical=False`` routes the uncached jobs through
  :func:`vardepth_amd.trie_vector.score` (the prefix-sharing path: values at
  or below 1e-200 are recomputed densely), ``canonical=True`` straight to the
  dense FP64 ``pairhmm_pairs``;
* results come back in job order.

How it is done here: the job list becomes a :class:`~vardepth_amd.jobs.JobTable`
(one integer triple per job), the distinct triples are found with one
``np.unique``, the cache is consulted once per distinct triple, everything
missing goes to the device in one call, and one fancy-indexing gather puts
the values back in job order.  The dense path hands ``pairhmm_pairs`` pointer
arrays into one packed buffer per string kind; its jobs are not reordered,
because the device plans its own work (the reference sorts by length to fill
its 8-lane SIMD groups, which changes no value).  ``lanes`` only takes part in
the cache key: the reference's lane count changes no result either.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .jobs import JobTable, Packed
from .native import cpu_batch_library as library  # noqa: F401  (reference name)

_CACHE_LIMIT = 131072


class _Recent:
    """Least-recently-used map: a dict kept in recency order (oldest first).

    A hit re-inserts its key at the young end; an insert beyond the module's
    ``_CACHE_LIMIT`` drops the oldest keys.  Iteration yields keys oldest first."""

    def __init__(self):
        self._d = {}

    def lookup(self, key):
        value = self._d.pop(key, None)
        if value is not None:
            self._d[key] = value
        return value

    def store(self, key, value: float):
        self._d.pop(key, None)
        self._d[key] = value
        while len(self._d) > _CACHE_LIMIT:
            del self._d[next(iter(self._d))]

    def clear(self):
        self._d.clear()

    def __len__(self):
        return len(self._d)

    def __iter__(self):
        return iter(list(self._d))

    def __contains__(self, key):
        return key in self._d


_CACHE = _Recent()


def pairs(jobs, mode=0, lanes=0, *, canonical=False):
    """(P,) ln likelihoods of (read, quality, haplotype) jobs (cpu_batch.py:20-42)."""
    jobs = list(jobs)
    if not jobs:
        return np.empty(0, np.float64)
    table = JobTable(jobs)
    first, inverse = table.unique()
    keys = []
    values = np.empty(len(first), np.float64)
    todo = []
    for u, j in enumerate(first):
        s, q, h = table.job(j)
        key = (mode, lanes, s, q, h, canonical)
        keys.append(key)
        hit = _CACHE.lookup(key)
        if hit is None:
            todo.append(u)
        else:
            values[u] = hit
    if todo:
        missing = [(k[2], k[3], k[4]) for k in (keys[u] for u in todo)]
        if canonical:
            computed = _uncached_pairs(missing, mode, lanes)
        else:
            from .trie_vector import score
            computed = score(missing, mode, _uncached_pairs, lanes=lanes)
        computed = np.asarray(computed, np.float64)
        values[todo] = computed
        for u, v in zip(todo, computed):
            _CACHE.store(keys[u], float(v))
    return values[inverse]


def _uncached_pairs(jobs, mode, lanes):
    """Dense FP64 likelihood of every job through the C-ABI ``pairhmm_pairs``
    (cpu_batch.py:44-62), one device call for the whole list."""
    count = len(jobs)
    table = JobTable(jobs)
    reads, quals, haps = Packed(table.seqs), Packed(table.quals), Packed(table.haps)
    rp = reads.pointers(table.seq)
    hp = haps.pointers(table.hap)
    qp = quals.pointers(table.qual)
    rn = reads.lengths[table.seq].astype(np.int32)
    hn = haps.lengths[table.hap].astype(np.int32)
    out = np.empty(max(count, 1), np.float64)
    ip = ctypes.POINTER(ctypes.c_int)
    code = library().pairhmm_pairs(rp, ctypes.cast(qp, ctypes.POINTER(ctypes.POINTER(ctypes.c_ubyte))),
                                   rn.ctypes.data_as(ip), hp, hn.ctypes.data_as(ip), count, mode, lanes,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if code:
        raise RuntimeError('HIP PairHMM failed: %d (%s)' % (code, _lib.last_error()))
    return out[:count]


def many(sequences, qualities, haplotypes, mode=0):
    """(R, H) matrix of R reads against H haplotypes (cpu_batch.py:64-66)."""
    if len(sequences) != len(qualities):
        raise ValueError('Read batch lengths differ')
    R, H = len(sequences), len(haplotypes)
    grid = [(s, q, h) for s, q in zip(sequences, qualities) for h in haplotypes]
    return pairs(grid, mode).reshape(R, H)
