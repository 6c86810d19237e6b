"""ctypes loaders with the reference's names and argtypes, bound to libvdhmm.so.

Each function returns the HIP library with exactly the prototypes the
reference loader it replaces declares, so code that calls
``library().pairhmm_pairs(...)`` and the like runs unchanged:

============================  ==========================================  ==========================
loader                        reference loader                            symbols
============================  ==========================================  ==========================
:func:`cpu_batch_library`     ``cpu_batch.library`` cpu_batch.py:11-18    ``pairhmm_pairs``
:func:`trie_library`          ``trie_vector.library`` trie_vector.py:8-14 ``trie_many``
:func:`core`                  ``base/caller.core`` base/caller.py:30-43   ``read_haplotype``,
                                                                          ``read_haplotype_batch``
                                                                          (mode 0, core semantics)
:func:`kernel`                ``indel_refine.kernel`` indel_refine.py:54  ``anchored_pairhmm``
:func:`wide_library`          ``whole_pairhmm.library``                   ``read_haplotype_batch``,
                              whole_pairhmm.py:16-48                      ``read_haplotype_batch_mode``,
                                                                          ``wide_capability``
============================  ==========================================  ==========================

``core()`` loads ``libvdhmm_core.so``, whose ``read_haplotype`` symbols carry
the core variant's semantics (mode 0, ``-inf`` when a row maximum is not
positive); ``libvdhmm.so``'s same-named symbols are the wide (mode 2) ones,
as in the reference where the two live in different libraries.
"""
from __future__ import annotations

import ctypes
import functools

from . import _lib

_u8p = ctypes.POINTER(ctypes.c_ubyte)
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


@functools.lru_cache(maxsize=1)
def cpu_batch_library() -> ctypes.CDLL:
    lib = _lib.lib()
    u = _u8p
    lib.pairhmm_pairs.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(u), _ip,
                                  ctypes.POINTER(ctypes.c_char_p), _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  _dp]
    lib.pairhmm_pairs.restype = ctypes.c_int
    return lib


@functools.lru_cache(maxsize=1)
def trie_library() -> ctypes.CDLL:
    lib = _lib.lib()
    p, i = ctypes.POINTER, ctypes.c_int
    lib.trie_many.argtypes = [p(ctypes.c_char_p), p(p(ctypes.c_ubyte)), p(i), i, p(ctypes.c_char_p), p(i), i, i, i,
                              p(ctypes.c_double), p(ctypes.c_uint64)]
    lib.trie_many.restype = i
    return lib


@functools.lru_cache(maxsize=1)
def core() -> ctypes.CDLL:
    if not _lib.CORE_LIB_PATH.exists():
        raise _lib.VdhmmError("libvdhmm_core.so is not built: run `python -m vardepth_amd.build`")
    _lib.lib()  # resolves libvdhmm.so (the core library's dependency) with a clear error
    lib = ctypes.CDLL(str(_lib.CORE_LIB_PATH))
    lib.read_haplotype.argtypes = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double]
    lib.read_haplotype.restype = ctypes.c_double
    lib.read_haplotype_batch.argtypes = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _ip,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
    lib.read_haplotype_batch.restype = None
    return lib


@functools.lru_cache(maxsize=1)
def kernel() -> ctypes.CDLL:
    lib = _lib.lib()
    lib.anchored_pairhmm.argtypes = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double]
    lib.anchored_pairhmm.restype = ctypes.c_double
    return lib


@functools.lru_cache(maxsize=1)
def wide_library() -> ctypes.CDLL:
    lib = _lib.lib()
    arguments = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _ip, ctypes.c_int,
                 ctypes.c_double, ctypes.c_double, _dp]
    lib.read_haplotype_batch_mode.argtypes = arguments + [ctypes.c_int]
    lib.read_haplotype_batch_mode.restype = None
    lib.read_haplotype_batch.argtypes = arguments
    lib.read_haplotype_batch.restype = None
    lib.read_haplotype.argtypes = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double]
    lib.read_haplotype.restype = ctypes.c_double
    lib.wide_capability.argtypes = []
    lib.wide_capability.restype = ctypes.c_int
    return lib
