"""Pair-list likelihoods on the MI355X: the interface of ``accelerator.py``.

Entry points with the reference's names, arguments and errors
(accelerator.py:25-62,80-190): :func:`validate_pairs`, :func:`cpu_pairs`,
:func:`cpu_simd_pairs`, :func:`score_pairs`, :func:`score_many`, and the
engine class :class:`HipPairHMM` (``CudaPairHMM`` is an alias) with a
``score`` method and a ``last`` telemetry dict.

* The engine groups the pair list by read (one region per distinct read and
  quality string, its haplotypes in pair order), sends every region of a
  memory-budget chunk to the device in one persistent launch, and rescues the
  pairs that leave the FP32 window in FP64 on the device.
* ``cpu_pairs`` and ``cpu_simd_pairs`` are the reference's exact FP64 scorers
  (scalar ``anchored_pairhmm`` / wide ``read_haplotype`` at g=.001, x=.1, and
  the wide batch over haplotypes sharing a read).  Here they run the same
  FP64 models on the device -- anchored mode 1 or semiglobal mode 2, every
  pair in FP64, one batched launch per call -- held to 1e-9 ln of the
  reference's CPU values.  The names stay so that callers which pass them as
  the canonical scorer (``scorer=cpu_pairs``) keep working.
* There is no CPU path behind any of them: without the HIP library or a
  device they raise.

Environment (read as the reference reads it, accelerator.py:82-90,161-173):
``VARDEPTH_PAIRHMM_BACKEND`` in {cpu, cuda, auto, hip} is validated (every
value runs on this engine); ``VARDEPTH_CUDA_PRECISION`` float64 (default) or
mixed; ``VARDEPTH_CUDA_DEVICE`` (default 0); ``VARDEPTH_CUDA_MEMORY_MB``
(default 256, at least 8): the read, quality, haplotype and output bytes of
one launch stay within it; ``VARDEPTH_ACCELERATOR_LOG``: ``last`` appended as
one JSON line per call.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import _lib, engine
from .jobs import JobTable

_ERROR_LUT = np.asarray([math.pow(10., -min(60, v) / 10.) for v in range(256)], dtype=np.float64)
_BACKENDS = ('cpu', 'cuda', 'auto', 'hip')
_DNA_OK = np.zeros(256, bool)
_DNA_OK[np.frombuffer(b"ACGTN", np.uint8)] = True


def _dna_ok(seq) -> bool:
    if not isinstance(seq, str):
        return False
    try:
        raw = seq.encode('ascii')
    except UnicodeEncodeError:
        return False
    return bool(_DNA_OK[np.frombuffer(raw, np.uint8)].all())


def _phred(q) -> np.ndarray:
    """uint8 copy of a quality string, or ValueError (integer Phred in [0, 255], 1-D)."""
    a = np.frombuffer(q, np.uint8) if isinstance(q, bytes) else np.asarray(q)
    bad = a.ndim != 1
    if not bad and a.size:
        kind = a.dtype.kind
        if kind in 'iub':
            bad = bool(a.min() < 0 or a.max() > 255)
        elif kind == 'f':
            bad = not bool(np.all(np.isfinite(a)) and np.all(a == np.floor(a)) and a.min() >= 0 and a.max() <= 255)
        else:
            bad = True
    if bad:
        raise ValueError('Base qualities must be integer Phred scores in [0,255]')
    return a.astype(np.uint8)


def validate_pairs(pairs):
    """The input contract (accelerator.py:25-40): reads and haplotypes are ``str``
    over ``ACGTN``, qualities bytes or a 1-D integer sequence in [0, 255] of the
    read's length.  Returns (s, uint8 array q, h) triples; raises ValueError
    with the reference's messages.  Each distinct string and quality object is
    checked once."""
    seen_dna = set()
    by_object = {}   # id(quality) -> (the object, its uint8 array); the object pins the id
    out = []
    for read, qual, hap in pairs:
        if len(read) != len(qual):
            raise ValueError('Read/quality lengths differ')
        for dna in (read, hap):
            if dna not in seen_dna:
                if not _dna_ok(dna):
                    raise ValueError('Uppercase DNA required')
                seen_dna.add(dna)
        entry = by_object.get(id(qual))
        if entry is None:
            entry = by_object[id(qual)] = (qual, _phred(qual))
        out.append((read, entry[1], hap))
    return out


def _read_regions(table: JobTable, jobs):
    """Regions of one read each: (representative job per region, the haplotype
    ids of its jobs in job order, the job indices in that order)."""
    keys = table.read_key()[jobs]
    order = np.argsort(keys, kind="stable")
    bounds = np.flatnonzero(np.diff(keys[order])) + 1
    groups = np.split(jobs[order], bounds) if len(jobs) else []
    groups.sort(key=lambda g: g[0])               # regions in order of their first job
    return [g[:1] for g in groups], [table.hap[g] for g in groups], groups


def _regions_by_read(pairs, order):
    """Group the pair indices ``order`` by (read, quality bytes): one region per
    distinct read, in order of first appearance.  Returns (regions, slots) with
    regions[k] = ([s], [q bytes], haplotypes) and slots[k] its pair indices."""
    order = np.asarray(list(order), np.int64)
    table = JobTable([pairs[i] for i in order], 'Read/quality lengths differ')
    reps, haps, groups = _read_regions(table, np.arange(len(order)))
    regions = [([table.seq_objects[table.seq[r[0]]]], [table.quals[table.qual[r[0]]]],
                [table.hap_objects[h] for h in hs]) for r, hs in zip(reps, haps)]
    return regions, [order[g].tolist() for g in groups]


def _fp64_scores(pairs, mode: int) -> np.ndarray:
    """Every pair in FP64 on the device at g=.001, x=.1 (empty -> -inf)."""
    pairs = validate_pairs(pairs)
    out = np.full(len(pairs), -np.inf)
    live = np.fromiter((bool(s) and bool(h) for s, _, h in pairs), bool, len(pairs))
    if not live.any():
        return out
    table = JobTable(pairs)
    reps, haps, groups = _read_regions(table, np.flatnonzero(live))
    b = table.region_batch(reps, haps, 0.001, 0.1)
    vals = engine.forward(b, mode=mode, precision="float64")
    out[np.concatenate(groups)] = vals
    return out


def cpu_pairs(pairs, anchored=False):
    """FP64 likelihood per pair (accelerator.py:42-53): anchored -> the global
    model of ``anchored_pairhmm`` (mode 1), else the semiglobal model of wide
    ``read_haplotype`` (mode 2); computed on the device, every pair in FP64."""
    return _fp64_scores(list(pairs), 1 if anchored else 2)


def cpu_simd_pairs(pairs):
    """Exact semiglobal FP64 likelihoods, haplotypes sharing a read batched
    together (accelerator.py:55-62); computed on the device."""
    return _fp64_scores(list(pairs), 2)


class HipPairHMM:
    """The MI355X engine behind :func:`score_pairs` (counterpart of
    ``CudaPairHMM``, accelerator.py:80-159)."""

    def __init__(self, device=0, memory_mb=256, precision=None, strategy=None):
        precision = precision or os.environ.get("VARDEPTH_CUDA_PRECISION", "float64")
        if precision not in ("float64", "mixed"):
            raise ValueError("CUDA precision must be float64 or mixed")
        if memory_mb < 8:
            raise ValueError('CUDA workspace budget must be at least 8 MiB')
        self.precision = precision
        self.device = int(device)
        self.memory_bytes = int(memory_mb) * 1024 ** 2
        self.strategy = strategy or "persistent"
        active = _lib.lib().vd_device()
        if active < 0:
            raise RuntimeError('HIP PairHMM requires an accessible MI355X GPU: ' + _lib.last_error())
        if active != self.device:
            raise ValueError('the HIP engine of this process runs on device %d, not %d' % (active, self.device))
        self.metadata = {"backend": "hip", "strategy": self.strategy, "device": self.device, "gpu": "gfx950",
                         "precision": precision, "fmad": True, "memory_budget_bytes": self.memory_bytes,
                         "library": str(_lib.LIB_PATH)}
        self.last = None

    def _chunks(self, need: np.ndarray, jobs: np.ndarray):
        """Split ``jobs`` (in order) into runs whose summed ``need`` bytes fit the budget."""
        if len(jobs) and need[jobs].max() > self.memory_bytes:
            raise ValueError('One pair exceeds CUDA workspace budget')
        start, used = 0, 0
        for i, j in enumerate(jobs):
            if used + need[j] > self.memory_bytes and i > start:
                yield jobs[start:i]
                start, used = i, 0
            used += need[j]
        if start < len(jobs):
            yield jobs[start:]

    def score(self, pairs, anchored=False):
        t_start = time.perf_counter()
        pairs = validate_pairs(pairs)
        result = np.full(len(pairs), -np.inf, dtype=np.float64)
        mode = 1 if anchored else 0
        stats = dict(kernel_ms=0.0, fwd64_ms=0.0, packing=0.0, batches=0, fp64_pairs=0, flagged=0, peak=0)
        if pairs:
            t0 = time.perf_counter()
            table = JobTable(pairs, 'Read/quality lengths differ')
            n = np.fromiter((len(b) for b in table.seqs), np.int64, len(table.seqs))[table.seq]
            m = np.fromiter((len(b) for b in table.haps), np.int64, len(table.haps))[table.hap]
            need = 2 * n + m + 64          # read + quality + haplotype bytes, offsets and output
            live = np.flatnonzero((n > 0) & (m > 0))
            stats["packing"] += time.perf_counter() - t0
            for chunk in self._chunks(need, live):
                t0 = time.perf_counter()
                reps, haps, groups = _read_regions(table, chunk)
                b = table.region_batch(reps, haps, 0.001, 0.1)
                stats["packing"] += time.perf_counter() - t0
                vals, st = engine.forward(b, mode=mode, precision=self.precision, stats=True)
                result[np.concatenate(groups)] = vals
                stats["kernel_ms"] += st["device_ms"]
                stats["fwd64_ms"] += st["fwd64_ms"]
                stats["fp64_pairs"] += st["fp64_pairs"]
                stats["flagged"] += st["flagged"]
                stats["batches"] += 1
                stats["peak"] = max(stats["peak"], int(2 * b.read_bytes.nbytes + b.hap_bytes.nbytes + 8 * b.npairs))
        mixed = self.precision == "mixed"
        self.last = dict(self.metadata)
        self.last.update(
            float64_fallback_backend='hip_float64_rescue' if mixed and stats["flagged"] else None,
            float64_fallback_pairs=int(stats["flagged"]) if mixed else 0,
            float64_fallback_seconds=stats["fwd64_ms"] / 1000. if mixed else 0.,
            float64_device_pairs=int(stats["fp64_pairs"]), pairs=len(pairs),
            dp_cells=int(sum(len(s) * len(h) for s, _, h in pairs)), batches=stats["batches"],
            live_array_bytes_upper_bound=stats["peak"], seconds=time.perf_counter() - t_start,
            includes_packing_transfers_and_synchronization=True, anchored=anchored,
            kernel_seconds=stats["kernel_ms"] / 1000., host_packing_seconds=stats["packing"])
        return result


CudaPairHMM = HipPairHMM   # the reference's class name


def backend():
    value = os.environ.get('VARDEPTH_PAIRHMM_BACKEND', 'cpu')
    if value not in _BACKENDS:
        raise ValueError('Invalid PairHMM backend')
    return value


_engine_cache = {}


def _engine() -> HipPairHMM:
    key = (int(os.environ.get('VARDEPTH_CUDA_DEVICE', '0')), int(os.environ.get('VARDEPTH_CUDA_MEMORY_MB', '256')),
           os.environ.get('VARDEPTH_CUDA_PRECISION', 'float64'))
    eng = _engine_cache.get(key)
    if eng is None:
        _engine_cache.clear()
        eng = _engine_cache[key] = HipPairHMM(*key)
    return eng


def _log_call(record):
    path = os.environ.get('VARDEPTH_ACCELERATOR_LOG')
    if not path:
        return
    import fcntl
    with open(path, 'a') as stream:
        fcntl.flock(stream.fileno(), fcntl.LOCK_EX)
        stream.write(json.dumps(record) + '\n')


def score_pairs(pairs, anchored=False, selection=None):
    """(P,) ln likelihoods of (read, quality, haplotype) triples; ``anchored``
    selects mode 1 (accelerator.py:166-186)."""
    if (selection or backend()) not in _BACKENDS:
        raise ValueError('Invalid PairHMM backend')
    pairs = list(pairs)
    if not pairs:
        return np.empty(0, np.float64)
    eng = _engine()
    values = eng.score(pairs, anchored)
    _log_call(eng.last)
    return values


def score_many(sequences, qualities, haplotypes, selection=None):
    """(R, H) matrix of R reads against H haplotypes (accelerator.py:188-190)."""
    if len(sequences) != len(qualities):
        raise ValueError('Sequence/quality batch lengths differ')
    grid = [(s, q, h) for s, q in zip(sequences, qualities) for h in haplotypes]
    return score_pairs(grid, selection=selection).reshape(len(sequences), len(haplotypes))
