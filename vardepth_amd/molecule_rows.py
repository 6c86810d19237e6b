"""Drop-in for the reference's ``molecule_rows`` (caller/native/caller/molecule_rows.py).

:func:`normalized_rows` keeps molecule_rows.py:6-14's arguments, validation
and outputs: values = (1-e) exp(ll - rowmax) + e mean with e =
10^(-min(mq,60)/10), and certainty = ptp(log(max(1e-300, values))) per row,
computed on the device (``vd_normalized_rows``).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_ubyte)


def normalized_rows(likelihoods, mapping_qualities):
    ll = np.asarray(likelihoods, dtype=np.float64)
    if ll.ndim != 2 or len(ll) != len(mapping_qualities):
        raise ValueError('Likelihood rows and MQ differ')
    if not len(ll):
        return ll.copy(), np.empty(0, np.float64)
    ll = np.ascontiguousarray(ll)
    mq = np.ascontiguousarray(np.asarray(mapping_qualities, np.uint8))
    rows, n = ll.shape
    if n == 0:
        raise ValueError('zero-size array to reduction operation maximum which has no identity')
    values = np.empty_like(ll)
    certainty = np.empty(rows, np.float64)
    _lib.check(_lib.lib().vd_normalized_rows(rows, n, ll.ctypes.data_as(_dp), mq.ctypes.data_as(_u8p),
                                             values.ctypes.data_as(_dp), certainty.ctypes.data_as(_dp)),
               "vd_normalized_rows")
    return values, certainty
