"""Region likelihood calls of the caller's base model, on the MI355X.

Mirrors the read x haplotype entry points of ``base/caller.py``
(base/caller.py:30-43,139-160) with the same names and arguments:

* :func:`core` -- the ctypes library with ``read_haplotype`` /
  ``read_haplotype_batch`` (core semantics: mode 0, parametric g, x,
  ``-inf`` when a row maximum is not positive);
* :func:`pair_score`, :func:`pair_scores` -- one read against one / several
  haplotypes at g=0.001, x=0.1 through that library;
* :func:`pair_scores_many` -- the region matrix (R, H) in mode 0, the call
  this engine replaces: ``VARDEPTH_PAIRHMM_BACKEND`` = cuda (or auto with
  sum(n) * sum(m) >= 2e6) goes to :func:`vardepth_amd.accelerator.score_many`,
  otherwise to :func:`vardepth_amd.batch.many`; both run on the GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .native import core  # noqa: F401  (reference name)


def ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def pair_score(lib, seq, quals, hap):
    qs = (ctypes.c_ubyte * len(quals))(*quals)
    return lib.read_haplotype(seq.encode(), qs, len(seq), hap.encode(), len(hap), .001, .1)


def pair_scores(lib, seq, quals, haps):
    qs = (ctypes.c_ubyte * len(quals))(*quals)
    encoded = [h.encode() for h in haps]
    hp = (ctypes.c_char_p * len(encoded))(*encoded)
    lengths = (ctypes.c_int * len(encoded))(*(len(h) for h in haps))
    out = np.empty(len(encoded), np.float64)
    lib.read_haplotype_batch(seq.encode(), qs, len(seq), hp, lengths, len(encoded), .001, .1, ptr(out))
    return out


def pair_scores_many(lib, sequences, qualities, haps):
    """(R, H) natural-log likelihoods of R reads against H haplotypes, mode 0.

    ``lib`` is accepted for signature compatibility (the reference passes its
    core library) and is not used."""
    choice = os.environ.get('VARDEPTH_PAIRHMM_BACKEND', 'cpu')
    cells = sum(map(len, sequences)) * sum(map(len, haps))
    if choice == 'cuda' or choice == 'hip' or (choice == 'auto' and cells >= 2_000_000):
        from .accelerator import score_many
        return score_many(sequences, qualities, haps)
    from .batch import many
    return many(sequences, qualities, haps, mode=0)
