"""ctypes binding of libvdhmm.so (the C ABI declared in include/vdhmm.h).

The library is built in-tree by ``python -m vardepth_amd.build``.  There is
no CPU fallback: if the library is missing or no HIP device is present, the
calls raise.
"""
from __future__ import annotations

import ctypes
import functools
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libvdhmm.so"
CORE_LIB_PATH = PKG / "libvdhmm_core.so"

_c_char_pp = ctypes.POINTER(ctypes.c_char_p)
_u8p = ctypes.POINTER(ctypes.c_ubyte)
_u8pp = ctypes.POINTER(_u8p)
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)

# every extern "C" symbol include/vdhmm.h declares
SYMBOLS = ("pairhmm_pairs", "trie_many", "read_haplotype", "read_haplotype_batch",
           "read_haplotype_batch_mode", "wide_capability", "anchored_pairhmm",
           "vd_core_read_haplotype", "vd_core_read_haplotype_batch", "vd_params_default",
           "vd_forward", "vd_forward_device", "vd_grad", "vd_grad_device", "vd_forward_plain", "vd_forward_refdesign",
           "vd_tuning_get",
           "vd_tuning_set", "vd_class_timing", "vd_class_ms", "vd_mixture_general", "vd_mixture_two",
           "vd_normalized_rows", "vd_mixture_regions_device", "vd_genotype_device", "vd_genotype_guard", "vd_prefix_sharing",
           "vd_suffix_sharing", "vd_class_overlap", "vd_max_hap_len",
           "vd_forward_device_noalloc", "vd_last_error", "vd_device")


class VdParams(ctypes.Structure):
    _fields_ = [("gap", ctypes.c_double), ("ext", ctypes.c_double), ("err", ctypes.c_double * 64)]


class VdBatch(ctypes.Structure):
    _fields_ = [("read_bytes", ctypes.c_void_p), ("qual_bytes", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("hap_bytes", ctypes.c_void_p),
                ("hap_off", ctypes.c_void_p), ("region_reads", ctypes.c_void_p),
                ("region_haps", ctypes.c_void_p), ("out_off", ctypes.c_void_p),
                ("nregions", ctypes.c_int64), ("nreads", ctypes.c_int64), ("nhaps", ctypes.c_int64)]


class VdStats(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_uint64), ("cells", ctypes.c_uint64), ("fp32_units", ctypes.c_uint64),
                ("flagged", ctypes.c_uint64), ("fp64_pairs", ctypes.c_uint64), ("fp64_bad", ctypes.c_uint64),
                ("overflow", ctypes.c_uint64), ("device_ms", ctypes.c_double), ("plan_ms", ctypes.c_double),
                ("fwd32_ms", ctypes.c_double), ("fwd64_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("fp32_cells_swept", ctypes.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class VdTuning(ctypes.Structure):
    _fields_ = [("fwd_unroll", ctypes.c_int * 17), ("fwd_blocks_per_cu", ctypes.c_int * 17),
                ("grad_variant", ctypes.c_int * 17), ("grad_blocks_per_cu", ctypes.c_int * 17),
                ("fwd_share", ctypes.c_int * 17), ("share_chunk", ctypes.c_int), ("class_overlap", ctypes.c_int),
                ("fwd_route64", ctypes.c_int * 17), ("fwd_pack", ctypes.c_int)]
    SCALARS = ("share_chunk", "class_overlap", "fwd_pack")

    def as_dict(self):
        return {name: (int(getattr(self, name)) if name in self.SCALARS else list(getattr(self, name)))
                for name, _ in self._fields_}


class VdGenoBatch(ctypes.Structure):
    _fields_ = [("row_off", ctypes.c_void_p), ("col_off", ctypes.c_void_p), ("mat_off", ctypes.c_void_p),
                ("gl_off", ctypes.c_void_p), ("nregions", ctypes.c_int64), ("nrows", ctypes.c_int64),
                ("mat_size", ctypes.c_int64), ("gl_size", ctypes.c_int64)]


class VdGenoStats(ctypes.Structure):
    _fields_ = [("units", ctypes.c_int64), ("kept_rows", ctypes.c_int64), ("invalid", ctypes.c_int64),
                ("skipped", ctypes.c_int64), ("molecule_ms", ctypes.c_double), ("mixture_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class VdhmmError(RuntimeError):
    pass


@functools.lru_cache(maxsize=1)
def lib() -> ctypes.CDLL:
    """Load libvdhmm.so and declare argtypes (raises if it has not been built)."""
    if not LIB_PATH.exists():
        raise VdhmmError("libvdhmm.so is not built: run `python -m vardepth_amd.build` "
                         "(the HIP engine has no CPU fallback)")
    # One HIP runtime per process: torch carries its own libamdhip64, and the
    # library's dependency on libamdhip64.so.7 binds to whichever copy is loaded
    # first.  Loading torch's first lets torch and the engine share it (loading
    # the system copy first leaves torch unable to see the device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(str(LIB_PATH))
    L.pairhmm_pairs.argtypes = [_c_char_pp, _u8pp, _ip, _c_char_pp, _ip, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, _dp]
    L.pairhmm_pairs.restype = ctypes.c_int
    L.trie_many.argtypes = [_c_char_pp, _u8pp, _ip, ctypes.c_int, _c_char_pp, _ip, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, _dp, _u64p]
    L.trie_many.restype = ctypes.c_int
    args7 = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_double,
             ctypes.c_double]
    batch = [ctypes.c_char_p, _u8p, ctypes.c_int, _c_char_pp, _ip, ctypes.c_int, ctypes.c_double,
             ctypes.c_double, _dp]
    for name in ("read_haplotype", "anchored_pairhmm", "vd_core_read_haplotype"):
        getattr(L, name).argtypes = args7
        getattr(L, name).restype = ctypes.c_double
    for name in ("read_haplotype_batch", "vd_core_read_haplotype_batch"):
        getattr(L, name).argtypes = batch
        getattr(L, name).restype = None
    L.read_haplotype_batch_mode.argtypes = batch + [ctypes.c_int]
    L.read_haplotype_batch_mode.restype = None
    L.wide_capability.argtypes = []
    L.wide_capability.restype = ctypes.c_int
    L.vd_params_default.argtypes = [ctypes.POINTER(VdParams), ctypes.c_double, ctypes.c_double]
    L.vd_params_default.restype = None
    L.vd_forward.argtypes = [ctypes.POINTER(VdBatch), ctypes.POINTER(VdParams), ctypes.c_int, ctypes.c_int,
                             _dp, ctypes.POINTER(VdStats)]
    L.vd_forward.restype = ctypes.c_int
    L.vd_forward_device.argtypes = [ctypes.POINTER(VdBatch), ctypes.POINTER(VdParams), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.POINTER(VdStats)]
    L.vd_forward_device.restype = ctypes.c_int
    L.vd_grad.argtypes = [ctypes.POINTER(VdBatch), ctypes.POINTER(VdParams), ctypes.c_int, _dp, _dp, _dp,
                          ctypes.POINTER(VdStats)]
    L.vd_grad.restype = ctypes.c_int
    L.vd_grad_device.argtypes = [ctypes.POINTER(VdBatch), ctypes.POINTER(VdParams), ctypes.c_int, ctypes.c_int64,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.POINTER(VdStats)]
    L.vd_grad_device.restype = ctypes.c_int
    L.vd_forward_plain.argtypes = [ctypes.POINTER(VdBatch), ctypes.POINTER(VdParams), ctypes.c_int, ctypes.c_int64,
                                   _dp, _dp]
    L.vd_forward_plain.restype = ctypes.c_int
    L.vd_forward_refdesign.argtypes = [ctypes.POINTER(VdBatch), ctypes.POINTER(VdParams), ctypes.c_int,
                                       ctypes.c_int64, _dp, _dp, _i64p]
    L.vd_forward_refdesign.restype = ctypes.c_int
    L.vd_tuning_get.argtypes = [ctypes.POINTER(VdTuning)]
    L.vd_tuning_get.restype = ctypes.c_int
    L.vd_tuning_set.argtypes = [ctypes.POINTER(VdTuning)]
    L.vd_tuning_set.restype = ctypes.c_int
    L.vd_prefix_sharing.argtypes = [ctypes.c_int]
    L.vd_prefix_sharing.restype = ctypes.c_int
    L.vd_suffix_sharing.argtypes = [ctypes.c_int]
    L.vd_suffix_sharing.restype = ctypes.c_int
    L.vd_class_overlap.argtypes = [ctypes.c_int]
    L.vd_class_overlap.restype = ctypes.c_int
    L.vd_class_timing.argtypes = [ctypes.c_int]
    L.vd_class_timing.restype = None
    L.vd_class_ms.argtypes = [_dp, _dp]
    L.vd_class_ms.restype = ctypes.c_int
    L.vd_mixture_general.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, _dp]
    L.vd_mixture_general.restype = ctypes.c_int
    L.vd_mixture_two.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp]
    L.vd_mixture_two.restype = ctypes.c_int
    L.vd_normalized_rows.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _u8p, _dp, _dp]
    L.vd_normalized_rows.restype = ctypes.c_int
    L.vd_mixture_regions_device.argtypes = [ctypes.POINTER(VdGenoBatch), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(VdGenoStats)]
    L.vd_mixture_regions_device.restype = ctypes.c_int
    L.vd_genotype_device.argtypes = [ctypes.POINTER(VdGenoBatch), ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(VdGenoStats)]
    L.vd_genotype_device.restype = ctypes.c_int
    L.vd_genotype_guard.argtypes = [ctypes.POINTER(VdGenoBatch), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
    L.vd_genotype_guard.restype = ctypes.c_int
    L.vd_forward_device_noalloc.argtypes = [ctypes.POINTER(VdBatch), ctypes.POINTER(VdParams), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p]
    L.vd_forward_device_noalloc.restype = ctypes.c_int
    L.vd_max_hap_len.argtypes = [ctypes.c_int]
    L.vd_max_hap_len.restype = ctypes.c_int
    L.vd_last_error.argtypes = []
    L.vd_last_error.restype = ctypes.c_char_p
    L.vd_device.argtypes = []
    L.vd_device.restype = ctypes.c_int
    return L


def last_error() -> str:
    return lib().vd_last_error().decode(errors="replace")


def check(code: int, what: str):
    if code:
        raise VdhmmError("%s failed (%d): %s" % (what, code, last_error()))


def params(gap: float, ext: float, err=None) -> VdParams:
    """vd_params with the reference error table (or a 64-bin / 256-entry table)."""
    p = VdParams()
    lib().vd_params_default(ctypes.byref(p), float(gap), float(ext))
    if err is not None:
        e = np.asarray(err, dtype=np.float64)
        if e.shape == (256,):
            e = e[:64]
        if e.shape != (64,):
            raise ValueError("error table must have 64 bins (or 256 raw-quality entries)")
        for i in range(64):
            p.err[i] = float(e[i])
    return p
