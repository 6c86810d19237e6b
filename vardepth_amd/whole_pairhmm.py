"""Prepared haplotype sets on the MI355X: the interface of ``whole_pairhmm.py``.

:class:`Prepared` has the reference's constructor, :meth:`~Prepared.many` and
``__call__`` (whole_pairhmm.py:51-81):

* the haplotypes are packed once into one buffer with a pointer array and
  lengths, held for the object's life;
* ``__call__(sequence, qualities)`` scores one read against every prepared
  haplotype through the C-ABI ``read_haplotype_batch_mode`` (the parametric
  semiglobal model of wide.cpp at g = 0.001, x = 0.1, FP64 on the device);
* ``many(sequences, qualities)`` scores a read batch as an (R, H) matrix:
  with ``VARDEPTH_PAIRHMM_BACKEND`` other than cpu through
  :func:`vardepth_amd.accelerator.score_many` (mode 0), else through
  :func:`vardepth_amd.batch.many` in mode 2 -- the two modes are the same
  model and differ only in rounding;
* ``mode`` is the reference's SIMD lane request (0, 2, 4 or 8): validated,
  passed on, and without effect on any value.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .jobs import Packed
from .native import wide_library as library

LANE_REQUESTS = (0, 2, 4, 8)     # wide.cpp:342-357


class Prepared:
    def __init__(self, haplotypes, mode=0):
        if mode not in LANE_REQUESTS:
            raise ValueError('Unsupported lane request')
        self.lib = library()
        self.haplotypes = list(haplotypes)
        self.encoded = [h.encode('ascii') for h in self.haplotypes]
        self.n = len(self.encoded)
        self._packed = Packed(self.encoded)
        self.hp = self._packed.pointers(np.arange(self.n))
        self._lengths = self._packed.lengths.astype(np.int32)
        self.lengths = self._lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        self.mode = mode

    def many(self, sequences, qualities):
        if len(sequences) != len(qualities):
            raise ValueError('Sequence/quality batch lengths differ')
        from .accelerator import backend
        if backend() != 'cpu':
            from .accelerator import score_many
            return score_many(sequences, qualities, self.haplotypes)
        from .batch import many
        return many(sequences, qualities, self.haplotypes, mode=2)

    def __call__(self, sequence, qualities):
        if len(sequence) != len(qualities):
            raise ValueError('Read and quality lengths disagree')
        q = np.ascontiguousarray(np.frombuffer(qualities, np.uint8) if isinstance(qualities, bytes)
                                 else np.asarray(qualities, dtype=np.uint8))
        out = np.empty(self.n, np.float64)
        self.lib.read_haplotype_batch_mode(
            sequence.encode('ascii'), q.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), len(sequence), self.hp,
            self.lengths, self.n, .001, .1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.mode)
        return out
