"""Drop-in for the reference's ``mixture_cpu`` (caller/native/caller/mixture_cpu.py).

Same functions, arguments, validation and outputs; the mixture runs on the
device through ``vd_mixture_two`` / ``vd_mixture_general`` (include/vdhmm.h).
The module name is kept so that callers switch by import path only.

* :func:`two`      -- mixture_cpu.py:12-15 -> mixture_two.cpp:9, 2 x 2 ordered genotypes.
* :func:`general`  -- mixture_cpu.py:23-28 -> mixture_general.cpp:9, triangular
  (b = 0..n-1, a = 0..b) or square a-major output.
* :func:`library` / :func:`general_library` -- the loaded C-ABI library, whose
  ``vd_mixture_two`` / ``vd_mixture_general`` take the reference's ctypes
  arguments (mixture_cpu.py:10, :20) and return a status code.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_dp = ctypes.POINTER(ctypes.c_double)


def library():
    return _lib.lib()


def general_library():
    return _lib.lib()


def _ptr(a):
    return a.ctypes.data_as(_dp)


def two(lh, prior):
    """mixture_cpu.py:12-15."""
    x = np.ascontiguousarray(lh, np.float64)
    out = np.empty((2, 2))
    if not np.isfinite(x).all() or np.any(x < 0) or not 0 <= prior < 1:
        raise ValueError('Invalid mixture likelihood/prior')
    if x.ndim != 2 or x.shape[1] != 2:
        raise ValueError('mixture_two takes rows of two likelihoods')
    _lib.check(_lib.lib().vd_mixture_two(len(x), _ptr(x), float(prior), _ptr(out)), "vd_mixture_two")
    return out


def general(lh, prior, triangular=False):
    """mixture_cpu.py:23-28."""
    x = np.ascontiguousarray(lh, np.float64)
    n = x.shape[1]
    if n < 1 or not np.isfinite(x).all() or np.any(x < 0) or (prior is not None and not 0 <= prior < 1):
        raise ValueError('Invalid mixture likelihood/prior')
    out = np.empty(n * (n + 1) // 2 if triangular else n * n)
    _lib.check(_lib.lib().vd_mixture_general(len(x), n, _ptr(x), -1.0 if prior is None else float(prior),
                                             int(bool(triangular)), _ptr(out)), "vd_mixture_general")
    return out if triangular else out.reshape(n, n)
