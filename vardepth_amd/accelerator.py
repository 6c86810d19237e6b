"""Pair-list likelihoods on the MI355X: the host interface of ``accelerator.py``.

Same entry points, argument meaning and errors as the reference module
(accelerator.py:25-62,80-190): :func:`validate_pairs`, :func:`score_pairs`,
:func:`score_many` and an engine class with a ``score`` method and a
``last`` telemetry dict.  The engine is the HIP region engine: the pair list
is grouped by read into regions (one read x its haplotypes), all of them go
to the device in one persistent launch, and pairs that underflow the FP32
window are rescued in FP64 on the device.  There is no CPU path behind these
functions; when the HIP library or device is missing they raise.

Environment (as read by the reference, accelerator.py:82-90,161-173):

* ``VARDEPTH_PAIRHMM_BACKEND`` in {cpu, cuda, auto, hip}: validated like the
  reference; every value runs on the HIP engine, because this library is the
  GPU path (the reference's CPU kernels are not part of it).
* ``VARDEPTH_CUDA_PRECISION`` in {float64, mixed} (default float64 as in the
  reference): float64 runs every pair in FP64 on the device; mixed runs FP32
  with FP64 rescue (error <= 1e-4 log10 per pair).
* ``VARDEPTH_CUDA_DEVICE`` (default 0), ``VARDEPTH_CUDA_MEMORY_MB``
  (default 256, minimum 8): pair batches are split so that one launch's
  read/haplotype bytes and outputs stay within the budget.
* ``VARDEPTH_ACCELERATOR_LOG``: append ``last`` as one JSON line per call.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import _lib, engine
from .workloads import from_regions

_ERROR_LUT = np.asarray([math.pow(10., -min(60, v) / 10.) for v in range(256)], dtype=np.float64)
_DNA = frozenset("ACGTN")
_ENGINE = None
_ENGINE_KEY = None


def validate_pairs(pairs):
    """Check the input contract (accelerator.py:25-40) and return (s, uint8 q, h) triples.

    Reads and haplotypes must be ``str`` over ``ACGTN``; qualities bytes or an
    integer array in [0, 255] of the read's length.  Raises ValueError with the
    reference's messages."""
    result = []
    checked_sequences = set()
    checked_qualities = {}
    for sequence, qualities, haplotype in pairs:
        if len(sequence) != len(qualities):
            raise ValueError('Read/quality lengths differ')
        for dna in (sequence, haplotype):
            if dna not in checked_sequences:
                if not isinstance(dna, str) or not set(dna) <= _DNA:
                    raise ValueError('Uppercase DNA required')
                checked_sequences.add(dna)
        key = id(qualities)
        if key not in checked_qualities:
            q = np.frombuffer(qualities, np.uint8) if isinstance(qualities, bytes) else np.asarray(qualities)
            bad = q.ndim != 1
            if not bad and q.size:
                if q.dtype.kind in "fc":
                    bad = bool(np.any(~np.isfinite(q)) or np.any(q != np.floor(q)))
                elif q.dtype.kind not in "iub":
                    bad = True
                if not bad:
                    bad = bool(np.any(q < 0) or np.any(q > 255))
            if bad:
                raise ValueError('Base qualities must be integer Phred scores in [0,255]')
            # keep the object alive with its converted copy: ids cannot be recycled meanwhile
            checked_qualities[key] = (qualities, np.asarray(q, dtype=np.uint8))
        result.append((sequence, checked_qualities[key][1], haplotype))
    return result


def _regions_by_read(pairs, order):
    """Group pair indices by (read, quality bytes): one region per distinct read.

    Returns (regions, slots) where regions is a list of ([s], [q], haps) and
    slots[k] lists the pair indices of region k in haplotype order."""
    groups = {}
    for i in order:
        s, q, h = pairs[i]
        groups.setdefault((s, q.tobytes()), []).append(i)
    regions, slots = [], []
    for (s, qb), ids in groups.items():
        regions.append(([s], [qb], [pairs[i][2] for i in ids]))
        slots.append(ids)
    return regions, slots


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
        self.metadata = dict(backend='hip', strategy=self.strategy, device=self.device, gpu='gfx950',
                             precision=precision, fmad=True, memory_budget_bytes=self.memory_bytes,
                             library=str(_lib.LIB_PATH))
        self.last = None

    def _chunks(self, pairs, order):
        """Split pair indices (in order) into launches within the memory budget."""
        chunk, used = [], 0
        for i in order:
            s, q, h = pairs[i]
            need = 2 * len(s) + len(h) + 64  # read + quality + haplotype bytes, offsets and output
            if need > self.memory_bytes:
                raise ValueError('One pair exceeds CUDA workspace budget')
            if used + need > self.memory_bytes and chunk:
                yield chunk
                chunk, used = [], 0
            chunk.append(i)
            used += need
        if chunk:
            yield chunk

    def score(self, pairs, anchored=False):
        start = time.perf_counter()
        pairs = validate_pairs(pairs)
        result = np.full(len(pairs), -np.inf, dtype=np.float64)
        live = [i for i, (s, q, h) in enumerate(pairs) if s and h]
        mode = 1 if anchored else 0
        kernel_ms = fp64_ms = 0.0
        packing = 0.0
        batches = fp64_pairs = flagged = 0
        peak = 0
        for ids in self._chunks(pairs, live):
            t0 = time.perf_counter()
            regions, slots = _regions_by_read(pairs, ids)
            batch = from_regions(regions, 0.001, 0.1)
            packing += time.perf_counter() - t0
            out, st = engine.forward(batch, mode=mode, precision=self.precision, stats=True)
            oo = batch.out_off()
            for k, idx in enumerate(slots):
                result[idx] = out[oo[k]:oo[k + 1]]
            kernel_ms += st["device_ms"]
            fp64_ms += st["fwd64_ms"]
            fp64_pairs += st["fp64_pairs"]
            flagged += st["flagged"]
            batches += 1
            peak = max(peak, int(batch.read_bytes.nbytes * 2 + batch.hap_bytes.nbytes + 8 * batch.npairs))
        mixed = self.precision == "mixed"
        self.last = dict(**self.metadata, float64_fallback_backend='hip_float64_rescue' if mixed and flagged else None,
                         float64_fallback_pairs=int(flagged) if mixed else 0,
                         float64_fallback_seconds=fp64_ms / 1000. if mixed else 0.,
                         float64_device_pairs=int(fp64_pairs),
                         pairs=len(pairs), dp_cells=sum(len(s) * len(h) for s, q, h in pairs),
                         batches=batches, live_array_bytes_upper_bound=peak,
                         seconds=time.perf_counter() - start,
                         includes_packing_transfers_and_synchronization=True, anchored=anchored,
                         kernel_seconds=kernel_ms / 1000., host_packing_seconds=packing)
        return result


def backend():
    value = os.environ.get('VARDEPTH_PAIRHMM_BACKEND', 'cpu')
    if value not in ('cpu', 'cuda', 'auto', 'hip'):
        raise ValueError('Invalid PairHMM backend')
    return value


def _engine():
    global _ENGINE, _ENGINE_KEY
    key = (int(os.environ.get('VARDEPTH_CUDA_DEVICE', '0')), int(os.environ.get('VARDEPTH_CUDA_MEMORY_MB', '256')),
           os.environ.get('VARDEPTH_CUDA_PRECISION', 'float64'))
    if _ENGINE is None or _ENGINE_KEY != key:
        _ENGINE = HipPairHMM(*key)
        _ENGINE_KEY = key
    return _ENGINE


def score_pairs(pairs, anchored=False, selection=None):
    """(P,) ln likelihoods of (read, quality, haplotype) triples; anchored selects mode 1
    (accelerator.py:166-186)."""
    choice = selection or backend()
    pairs = list(pairs)
    if choice not in ('cpu', 'cuda', 'auto', 'hip'):
        raise ValueError('Invalid PairHMM backend')
    if not pairs:
        return np.empty(0, np.float64)
    eng = _engine()
    result = eng.score(pairs, anchored)
    telemetry = os.environ.get('VARDEPTH_ACCELERATOR_LOG')
    if telemetry:
        import fcntl
        with open(telemetry, 'a') as stream:
            fcntl.flock(stream.fileno(), fcntl.LOCK_EX)
            stream.write(json.dumps(eng.last) + '\n')
    return result


def score_many(sequences, qualities, haplotypes, selection=None):
    """(R, H) matrix for R reads x H haplotypes (accelerator.py:188-190)."""
    if len(sequences) != len(qualities):
        raise ValueError('Sequence/quality batch lengths differ')
    return score_pairs([(s, q, h) for s, q in zip(sequences, qualities) for h in haplotypes],
                       selection=selection).reshape(len(sequences), len(haplotypes))
