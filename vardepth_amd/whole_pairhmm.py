"""Prepared haplotype sets on the MI355X: the host interface of ``whole_pairhmm.py``.

:class:`Prepared` keeps the reference's constructor, ``many`` and ``__call__``
(whole_pairhmm.py:51-81): ``__call__`` scores one read against every prepared
haplotype through the C-ABI ``read_haplotype_batch_mode`` (parametric mode-2
semiglobal at g=0.001, x=0.1, one device launch per call); ``many`` scores a
read batch as an (R, H) matrix through :mod:`vardepth_amd.accelerator` or
:mod:`vardepth_amd.batch` (mode 2), selected by ``VARDEPTH_PAIRHMM_BACKEND``
as in the reference.  ``mode`` is the reference's SIMD lane request
(0, 2, 4 or 8); it is validated and does not change any value.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .native import wide_library as library


class Prepared:
    def __init__(self, haplotypes, mode=0):
        if mode not in [0, 2, 4, 8]:
            raise ValueError('Unsupported lane request')
        self.lib = library()
        self.encoded = [h.encode('ascii') for h in haplotypes]
        self.n = len(self.encoded)
        self.hp = (ctypes.c_char_p * self.n)(*self.encoded)
        self.lengths = (ctypes.c_int * self.n)(*(len(h) for h in self.encoded))
        self.mode = mode

    def many(self, sequences, qualities):
        if len(sequences) != len(qualities):
            raise ValueError('Sequence/quality batch lengths differ')
        from .accelerator import backend, score_many
        if backend() != 'cpu':
            return score_many(sequences, qualities, [h.decode('ascii') for h in self.encoded])
        from .batch import many
        return many(sequences, qualities, [h.decode('ascii') for h in self.encoded], mode=2)

    def __call__(self, sequence, qualities):
        if len(sequence) != len(qualities):
            raise ValueError('Read and quality lengths disagree')
        quality = np.frombuffer(qualities, np.uint8) if isinstance(qualities, bytes) else np.asarray(qualities,
                                                                                                   dtype=np.uint8)
        if not quality.flags.c_contiguous:
            quality = np.ascontiguousarray(quality)
        result = np.empty(self.n, np.float64)
        self.lib.read_haplotype_batch_mode(
            sequence.encode('ascii'), quality.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
            len(sequence), self.hp, self.lengths, self.n, .001, .1,
            result.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.mode)
        return result
