"""Cached pair batches on the MI355X: the host interface of ``cpu_batch.py``.

:func:`pairs`, :func:`many` and :func:`_uncached_pairs` keep the reference's
names, arguments and behaviour (cpu_batch.py:8-66):

* a likelihood is a pure function of (sequence, quality, haplotype, mode), so
  results are cached under ``(mode, lanes, s, q, h, canonical)`` in an LRU of
  131,072 entries and duplicate jobs in one call are computed once;
* ``canonical=False`` goes through :func:`vardepth_amd.trie_vector.score`
  (the prefix-sharing path's routing: values at or below 1e-200 are
  recomputed by the dense path), ``canonical=True`` straight to the dense
  FP64 ``pairhmm_pairs``;
* output order is the input order.

Both paths run on the GPU: the dense path is the C-ABI ``pairhmm_pairs`` of
libvdhmm.so (every pair in FP64 on the device), the trie path the batched
region engine.  ``lanes`` only takes part in the cache key, as the reference's
SIMD lane count does not change any result.
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict

import numpy as np

from . import _lib
from .native import cpu_batch_library as library  # noqa: F401  (reference name)

_CACHE = OrderedDict()
_CACHE_LIMIT = 131072


def _quality_bytes(q):
    return q if isinstance(q, bytes) else np.ascontiguousarray(q, dtype=np.uint8).tobytes()


def pairs(jobs, mode=0, lanes=0, *, canonical=False):
    """(P,) ln likelihoods of (read, quality, haplotype) jobs (cpu_batch.py:20-42)."""
    jobs = list(jobs)
    count = len(jobs)
    if not count:
        return np.empty(0, np.float64)
    quality = {}
    unique = {}
    missing = []
    slots = []
    result = np.empty(count, np.float64)
    for i, (s, q, h) in enumerate(jobs):
        if len(s) != len(q):
            raise ValueError('Read and quality lengths differ')
        if id(q) not in quality:
            quality[id(q)] = _quality_bytes(q)
        key = (mode, lanes, s, quality[id(q)], h, canonical)
        if key in _CACHE:
            result[i] = _CACHE[key]
            _CACHE.move_to_end(key)
        else:
            if key not in unique:
                unique[key] = len(missing)
                missing.append((s, key[3], h))
            slots.append((i, unique[key]))
    if not missing:
        return result
    from .trie_vector import score
    values = _uncached_pairs(missing, mode, lanes) if canonical else score(missing, mode, _uncached_pairs, lanes=lanes)
    for key, index in unique.items():
        _CACHE[key] = float(values[index])
        if len(_CACHE) > _CACHE_LIMIT:
            _CACHE.popitem(last=False)
    for i, j in slots:
        result[i] = values[j]
    return result


def _uncached_pairs(jobs, mode, lanes):
    """Dense FP64 likelihoods of every job through the C-ABI ``pairhmm_pairs``
    (cpu_batch.py:44-62): jobs sorted by (n, m), original order restored."""
    count = len(jobs)
    order = sorted(range(count), key=lambda j: (len(jobs[j][0]), len(jobs[j][2])))
    encoded = {}
    for j in jobs:
        for s in (j[0], j[2]):
            if s not in encoded:
                encoded[s] = s.encode('ascii') if isinstance(s, str) else bytes(s)
    quality = {}
    for j in jobs:
        if len(j[0]) != len(j[1]):
            raise ValueError('Read and quality lengths differ')
        if id(j[1]) not in quality:
            quality[id(j[1])] = _quality_bytes(j[1])
    reads = (ctypes.c_char_p * count)(*(encoded[jobs[i][0]] for i in order))
    haps = (ctypes.c_char_p * count)(*(encoded[jobs[i][2]] for i in order))
    u = ctypes.POINTER(ctypes.c_ubyte)
    qp = (ctypes.c_char_p * count)(*(quality[id(jobs[i][1])] for i in order))
    qs = ctypes.cast(qp, ctypes.POINTER(u))
    rn = np.array([len(jobs[i][0]) for i in order], np.int32)
    hn = np.array([len(jobs[i][2]) for i in order], np.int32)
    out = np.empty(count, np.float64)

    def ptr(a, t):
        return a.ctypes.data_as(ctypes.POINTER(t))

    code = library().pairhmm_pairs(reads, qs, ptr(rn, ctypes.c_int), haps, ptr(hn, ctypes.c_int), count, mode,
                                   lanes, ptr(out, ctypes.c_double))
    if code:
        raise RuntimeError('HIP PairHMM failed: %d (%s)' % (code, _lib.last_error()))
    result = np.empty_like(out)
    result[order] = out
    return result


def many(sequences, qualities, haplotypes, mode=0):
    """(R, H) matrix (cpu_batch.py:64-66)."""
    if len(sequences) != len(qualities):
        raise ValueError('Read batch lengths differ')
    return pairs([(s, q, h) for s, q in zip(sequences, qualities) for h in haplotypes],
                 mode).reshape(len(sequences), len(haplotypes))
