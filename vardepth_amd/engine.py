"""Batched region forward on the MI355X (host side of the region API).

:func:`forward` takes a :class:`~vardepth_amd.workloads.RegionBatch` in host
memory and returns the natural-log likelihood of every read x haplotype pair
in the batch's output layout.  :class:`DeviceBatch` keeps a batch resident in
HBM (torch tensors are only the allocator) so that repeated forwards measure
the device pipeline alone.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .workloads import RegionBatch

MODES = {0: 0, 1: 1, 2: 2, "semiglobal": 0, "anchored": 1, "rescale": 2}
PRECISIONS = {"mixed": 0, "float64": 1, "fp64": 1}


def _mode(mode) -> int:
    if mode not in MODES:
        raise ValueError("mode must be 0, 1 or 2")
    return MODES[mode]


def _precision(precision) -> int:
    if precision not in PRECISIONS:
        raise ValueError("precision must be 'mixed' or 'float64'")
    return PRECISIONS[precision]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def forward(batch: RegionBatch, mode=0, precision="mixed", err=None, stats=False):
    """ln likelihoods (float64, batch output layout); with stats=True also a stats dict."""
    oo = batch.out_off()
    keep = [batch.read_bytes, batch.qual_bytes, batch.read_off, batch.hap_bytes, batch.hap_off,
            batch.region_reads, batch.region_haps, oo]
    # non-empty buffers even for empty batches (the C ABI dereferences offsets[0])
    keep = [a if a.size else np.zeros(1, a.dtype) for a in keep]
    vb = _lib.VdBatch(*[_ptr(a) for a in keep], batch.nregions, batch.nreads, batch.nhaps)
    out = np.empty(max(batch.npairs, 1), np.float64)
    st = _lib.VdStats()
    p = _lib.params(batch.gap, batch.ext, err)
    code = _lib.lib().vd_forward(ctypes.byref(vb), ctypes.byref(p), _mode(mode), _precision(precision),
                                 out.ctypes.data_as(_lib._dp), ctypes.byref(st))
    _lib.check(code, "vd_forward")
    out = out[:batch.npairs]
    return (out, st.as_dict()) if stats else out


class DeviceBatch:
    """A RegionBatch copied once into HBM; forward() runs the device pipeline only."""

    def __init__(self, batch: RegionBatch, device="cuda:0"):
        import torch
        self.torch = torch
        self.batch = batch
        oo = batch.out_off()
        self.npairs = int(batch.npairs)
        self.max_hap = int(batch.hap_lengths().max()) if batch.nhaps else 0
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a if a.size else np.zeros(1, a.dtype))).to(device)
        self.tensors = [t(batch.read_bytes), t(batch.qual_bytes), t(batch.read_off), t(batch.hap_bytes),
                        t(batch.hap_off), t(batch.region_reads), t(batch.region_haps), t(oo)]
        self.out = torch.empty(max(self.npairs, 1), dtype=torch.float64, device=device)
        self.vb = _lib.VdBatch(*[x.data_ptr() for x in self.tensors], batch.nregions, batch.nreads, batch.nhaps)

    def forward(self, mode=0, precision="mixed", err=None, stream=None, stats=False):
        """Enqueue the forward on `stream` (torch stream or None = torch's current stream)."""
        torch = self.torch
        s = stream if stream is not None else torch.cuda.current_stream()
        p = _lib.params(self.batch.gap, self.batch.ext, err)
        st = _lib.VdStats() if stats else None
        code = _lib.lib().vd_forward_device(ctypes.byref(self.vb), ctypes.byref(p), _mode(mode),
                                            _precision(precision), self.npairs, self.max_hap,
                                            ctypes.c_void_p(self.out.data_ptr()),
                                            ctypes.c_void_p(s.cuda_stream),
                                            ctypes.byref(st) if st is not None else None)
        _lib.check(code, "vd_forward_device")
        return st.as_dict() if st is not None else None

    def result(self) -> np.ndarray:
        return self.out[:self.npairs].cpu().numpy()
