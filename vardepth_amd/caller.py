"""Region likelihood calls of the caller's base model, on the MI355X.

The read x haplotype entry points of ``base/caller.py`` (base/caller.py:30-43,
139-160), with the same names and arguments:

* :func:`core` -- the ctypes library exporting ``read_haplotype`` /
  ``read_haplotype_batch`` with the core variant's semantics (mode 0,
  parametric g and x, ``-inf`` when a row maximum is not positive);
* :func:`pair_score`, :func:`pair_scores` -- one read against one or several
  haplotypes at the base model's g = 0.001, x = 0.1, through that library;
* :func:`pair_scores_many` -- the region matrix (R, H) in mode 0, the call
  this engine replaces.  ``VARDEPTH_PAIRHMM_BACKEND`` = cuda or hip, or auto
  with sum(n) * sum(m) >= 2e6 cells, selects the pair-list engine
  (:func:`vardepth_amd.accelerator.score_many`); otherwise the cached batch
  path (:func:`vardepth_amd.batch.many`).  Both run on the GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .jobs import Packed, seq_bytes
from .native import core  # noqa: F401  (reference name)

BASE_GAP, BASE_EXT = .001, .1       # base/caller.py:139,148
AUTO_CELLS = 2_000_000               # base/caller.py:156


def ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _read_args(seq, quals):
    """(read bytes, uint8 quality buffer, n) for the read_haplotype calls."""
    raw = np.frombuffer(quals, np.uint8) if isinstance(quals, (bytes, bytearray)) else np.asarray(quals)
    q = (ctypes.c_ubyte * len(raw)).from_buffer_copy(raw.astype(np.uint8).tobytes())
    return seq_bytes(seq), q, len(seq)


def pair_score(lib, seq, quals, hap):
    """ln L of one read against one haplotype (base/caller.py:138-140)."""
    read, q, n = _read_args(seq, quals)
    h = seq_bytes(hap)
    return lib.read_haplotype(read, q, n, h, len(h), BASE_GAP, BASE_EXT)


def pair_scores(lib, seq, quals, haps):
    """ln L of one read against each haplotype (base/caller.py:142-149)."""
    read, q, n = _read_args(seq, quals)
    packed = Packed([seq_bytes(h) for h in haps])
    lengths = packed.lengths.astype(np.int32)
    out = np.empty(len(haps), np.float64)
    lib.read_haplotype_batch(read, q, n, packed.pointers(np.arange(len(haps))),
                             lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(haps), BASE_GAP, BASE_EXT,
                             ptr(out))
    return out


def _wants_engine(sequences, haps) -> bool:
    choice = os.environ.get('VARDEPTH_PAIRHMM_BACKEND', 'cpu')
    if choice in ('cuda', 'hip'):
        return True
    return choice == 'auto' and sum(map(len, sequences)) * sum(map(len, haps)) >= AUTO_CELLS


def pair_scores_many(lib, sequences, qualities, haps):
    """(R, H) natural-log likelihoods of R reads against H haplotypes, mode 0
    (base/caller.py:151-160).  ``lib`` is accepted for the reference's
    signature (it passes its core library) and is not used."""
    if _wants_engine(sequences, haps):
        from .accelerator import score_many
        return score_many(sequences, qualities, haps)
    from .batch import many
    return many(sequences, qualities, haps, mode=0)
