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


def grad(batch: RegionBatch, mode=0, err=None, weights=None, stats=False):
    """Gradient of sum_p w_p ln L_p (adjoint kernels).

    Returns (ln L per pair, grad) where grad is float64[66]: d/dgap, d/dext,
    then d/derr[q] for the 64 quality bins; with stats=True also a stats dict.
    ``weights``: one per output slot (None = all 1)."""
    oo = batch.out_off()
    keep = [batch.read_bytes, batch.qual_bytes, batch.read_off, batch.hap_bytes, batch.hap_off,
            batch.region_reads, batch.region_haps, oo]
    keep = [a if a.size else np.zeros(1, a.dtype) for a in keep]
    vb = _lib.VdBatch(*[_ptr(a) for a in keep], batch.nregions, batch.nreads, batch.nhaps)
    out = np.empty(max(batch.npairs, 1), np.float64)
    g = np.zeros(GRAD_SIZE, np.float64)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if w.shape != (batch.npairs,):
            raise ValueError("weights must have one entry per read x haplotype pair")
    st = _lib.VdStats()
    p = _lib.params(batch.gap, batch.ext, err)
    code = _lib.lib().vd_grad(ctypes.byref(vb), ctypes.byref(p), _mode(mode),
                              w.ctypes.data_as(_lib._dp) if w is not None else None,
                              out.ctypes.data_as(_lib._dp), g.ctypes.data_as(_lib._dp), ctypes.byref(st))
    _lib.check(code, "vd_grad")
    out = out[:batch.npairs]
    return (out, g, st.as_dict()) if stats else (out, g)


GRAD_SIZE = 66  # d/dgap, d/dext, d/derr[0..63]


def forward_plain(batch: RegionBatch, mode=0, err=None, max_pairs=None):
    """The one-pair-per-launch baseline on the first max_pairs pairs (region
    order): returns (ln L with NaN for pairs not run, device milliseconds)."""
    oo = batch.out_off()
    keep = [batch.read_bytes, batch.qual_bytes, batch.read_off, batch.hap_bytes, batch.hap_off,
            batch.region_reads, batch.region_haps, oo]
    keep = [a if a.size else np.zeros(1, a.dtype) for a in keep]
    vb = _lib.VdBatch(*[_ptr(a) for a in keep], batch.nregions, batch.nreads, batch.nhaps)
    out = np.full(max(batch.npairs, 1), np.nan)
    ms = ctypes.c_double(0.0)
    p = _lib.params(batch.gap, batch.ext, err)
    todo = batch.npairs if max_pairs is None else int(max_pairs)
    code = _lib.lib().vd_forward_plain(ctypes.byref(vb), ctypes.byref(p), _mode(mode), todo,
                                       out.ctypes.data_as(_lib._dp), ctypes.byref(ms))
    _lib.check(code, "vd_forward_plain")
    return out[:batch.npairs], ms.value


def forward_refdesign(batch: RegionBatch, mode=0, err=None, budget_mb=1024):
    """The reference's own GPU design (one thread per pair, DP rows in global
    memory, FP32 with its 2^-40 rescale rule), batched in length buckets, as a
    measurement baseline: returns (ln L with NaN where that design would fall
    back to FP64, device milliseconds, launches)."""
    oo = batch.out_off()
    keep = [batch.read_bytes, batch.qual_bytes, batch.read_off, batch.hap_bytes, batch.hap_off,
            batch.region_reads, batch.region_haps, oo]
    keep = [a if a.size else np.zeros(1, a.dtype) for a in keep]
    vb = _lib.VdBatch(*[_ptr(a) for a in keep], batch.nregions, batch.nreads, batch.nhaps)
    out = np.full(max(batch.npairs, 1), np.nan)
    ms = ctypes.c_double(0.0)
    launches = ctypes.c_int64(0)
    p = _lib.params(batch.gap, batch.ext, err)
    code = _lib.lib().vd_forward_refdesign(ctypes.byref(vb), ctypes.byref(p), _mode(mode), int(budget_mb),
                                           out.ctypes.data_as(_lib._dp), ctypes.byref(ms), ctypes.byref(launches))
    _lib.check(code, "vd_forward_refdesign")
    return out[:batch.npairs], ms.value, int(launches.value)


GENO_MIN_CERTAINTY = float(np.log(3))   # whole_read.py:37
GENO_PRIOR = 1e-3                        # whole_read.py:49
GUARD_MARGIN = 1e-4                      # partial_anchor.py:125: 1e-4 * max(1, molecules)


def gl_offsets(ncols, triangular=True) -> np.ndarray:
    """Pair-output offsets of regions with ``ncols`` haplotypes (length nregions+1)."""
    n = np.asarray(ncols, np.int64)
    pairs = n * (n + 1) // 2 if triangular else n * n
    return np.concatenate([[0], np.cumsum(pairs)]).astype(np.int64)


class DeviceBatch:
    """A RegionBatch copied once into HBM; forward() runs the device pipeline only."""

    def __init__(self, batch: RegionBatch, device="cuda:0"):
        import torch
        self.torch = torch
        self.batch = batch
        oo = batch.out_off()
        self.npairs = int(batch.npairs)
        self.max_hap = int(batch.hap_lengths().max()) if batch.nhaps else 0
        self.max_read = int(batch.read_lengths().max()) if batch.nreads else 0
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a if a.size else np.zeros(1, a.dtype))).to(device)
        self.tensors = [t(batch.read_bytes), t(batch.qual_bytes), t(batch.read_off), t(batch.hap_bytes),
                        t(batch.hap_off), t(batch.region_reads), t(batch.region_haps), t(oo)]
        self.out = torch.empty(max(self.npairs, 1), dtype=torch.float64, device=device)
        self.vb = _lib.VdBatch(*[x.data_ptr() for x in self.tensors], batch.nregions, batch.nreads, batch.nhaps)

    def forward(self, mode=0, precision="mixed", err=None, stream=None, stats=False, gap=None, ext=None):
        """Enqueue the forward on `stream` (torch stream or None = torch's current stream)."""
        torch = self.torch
        s = stream if stream is not None else torch.cuda.current_stream()
        p = _lib.params(self.batch.gap if gap is None else gap, self.batch.ext if ext is None else ext, err)
        st = _lib.VdStats() if stats else None
        code = _lib.lib().vd_forward_device(ctypes.byref(self.vb), ctypes.byref(p), _mode(mode),
                                            _precision(precision), self.npairs, self.max_hap,
                                            ctypes.c_void_p(self.out.data_ptr()),
                                            ctypes.c_void_p(s.cuda_stream),
                                            ctypes.byref(st) if st is not None else None)
        _lib.check(code, "vd_forward_device")
        return st.as_dict() if st is not None else None

    def graph(self, mode=0, precision="mixed", err=None, gap=None, ext=None):
        """Capture one forward (planning, FP32 kernel, FP64 rescue) in a HIP graph.

        A first eager forward sizes the device buffers and uploads the
        parameter tables, so the captured call only enqueues kernels (counters
        are cleared by a kernel, not a memset).  The graph keeps those buffers
        and tables: later calls of any size or parameters allocate their own,
        so replaying the returned torch.cuda.CUDAGraph recomputes self.out for
        the captured parameters at any later time (replays of one graph must
        not overlap)."""
        torch = self.torch
        self.forward(mode=mode, precision=precision, err=err, gap=gap, ext=ext)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        with torch.cuda.graph(g, stream=s):
            self.forward(mode=mode, precision=precision, err=err, gap=gap, ext=ext, stream=s)
        return g

    def grad(self, gap=None, ext=None, err=None, mode=0, weights=None, stream=None, stats=False):
        """Enqueue the adjoint pass; returns the float64[66] gradient tensor on the
        device (d/dgap, d/dext, d/derr[64]); ln L per pair lands in self.out.
        ``weights`` is a float64 device tensor with one entry per pair (or None)."""
        torch = self.torch
        s = stream if stream is not None else torch.cuda.current_stream()
        p = _lib.params(self.batch.gap if gap is None else gap, self.batch.ext if ext is None else ext, err)
        g = torch.zeros(GRAD_SIZE, dtype=torch.float64, device=self.out.device)
        w = None
        if weights is not None:
            w = weights.detach().to(device=self.out.device, dtype=torch.float64).contiguous()
            if w.numel() != self.npairs:
                raise ValueError("weights must have one entry per read x haplotype pair")
        st = _lib.VdStats() if stats else None
        code = _lib.lib().vd_grad_device(ctypes.byref(self.vb), ctypes.byref(p), _mode(mode), self.npairs,
                                         self.max_hap, self.max_read, ctypes.c_void_p(w.data_ptr()) if w is not None
                                         else None, ctypes.c_void_p(self.out.data_ptr()),
                                         ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(s.cuda_stream),
                                         ctypes.byref(st) if st is not None else None)
        _lib.check(code, "vd_grad_device")
        return (g, st.as_dict()) if st is not None else g

    def genotype(self, mq, min_certainty=GENO_MIN_CERTAINTY, prior=GENO_PRIOR, stream=None, stats=False,
                 values=False):
        """Enqueue the single-end genotype chain of whole_read.py:30-49 on the
        likelihoods in self.out (run forward() first): normalized rows with the
        per-read mapping qualities ``mq`` (uint8 device tensor, one per read),
        the certainty filter, molecule rows and the triangular diploid
        likelihood.  Returns the float64 device tensor of every region's pairs
        (region k at gl_off[k], b = 0..n-1, a = 0..b), plus (values, certainty)
        device tensors when ``values`` and the stats dict when ``stats``."""
        torch = self.torch
        s = stream if stream is not None else torch.cuda.current_stream()
        if mq.numel() != self.batch.nreads:
            raise ValueError("mq must hold one mapping quality per read")
        mq = mq.to(device=self.out.device, dtype=torch.uint8).contiguous()
        if not hasattr(self, "_gl_off"):
            off = gl_offsets(np.diff(self.batch.region_haps))
            self.gl_size = int(off[-1])
            self._gl_off = torch.from_numpy(off).to(self.out.device)
            self.gb = _lib.VdGenoBatch(self.tensors[5].data_ptr(), self.tensors[6].data_ptr(),
                                       self.tensors[7].data_ptr(), self._gl_off.data_ptr(), self.batch.nregions,
                                       self.batch.nreads, self.npairs, self.gl_size)
        gl = torch.empty(max(self.gl_size, 1), dtype=torch.float64, device=self.out.device)
        vals = cert = None
        if values:
            vals = torch.empty(max(self.npairs, 1), dtype=torch.float64, device=self.out.device)
            cert = torch.empty(max(self.batch.nreads, 1), dtype=torch.float64, device=self.out.device)
        st = _lib.VdGenoStats() if stats else None
        code = _lib.lib().vd_genotype_device(ctypes.byref(self.gb), ctypes.c_void_p(self.out.data_ptr()),
                                             ctypes.c_void_p(mq.data_ptr()), float(min_certainty), float(prior),
                                             ctypes.c_void_p(gl.data_ptr()),
                                             ctypes.c_void_p(vals.data_ptr()) if values else None,
                                             ctypes.c_void_p(cert.data_ptr()) if values else None,
                                             ctypes.c_void_p(s.cuda_stream),
                                             ctypes.byref(st) if st is not None else None)
        _lib.check(code, "vd_genotype_device")
        res = [gl[:self.gl_size]]
        if values:
            res += [vals[:self.npairs], cert[:self.batch.nreads]]
        if stats:
            res.append(st.as_dict())
        return res[0] if len(res) == 1 else tuple(res)

    def genotype_guarded(self, mq, mode=0, min_certainty=GENO_MIN_CERTAINTY, prior=GENO_PRIOR,
                         margin_scale=GUARD_MARGIN, stream=None, allele_freq=None):
        """The genotype chain with the near-tie guard of partial_anchor.py:119-133.

        After :meth:`genotype` on the mixed-precision likelihoods in self.out
        (run ``forward(mode=mode)`` first), every region whose two best
        genotypes -- by likelihood, or by likelihood plus the haplotype prior
        -- are within ``margin_scale * max(1, molecules)`` is recomputed from
        an all-FP64 forward of its reads and haplotypes, and its outputs
        replaced.  The prior comes from ``allele_freq`` (float64, one per
        haplotype of the batch: partial_anchor.py:56-57's panel frequencies,
        floored at .001) or, when None, is whole_read.py:51's uniform one.
        Returns (gl device tensor, report): the report lists the flagged
        regions with their observed gaps and margins."""
        torch = self.torch
        s = stream if stream is not None else torch.cuda.current_stream()
        gl, _, cert = self.genotype(mq, min_certainty, prior, stream=s, values=True)
        nreg = self.batch.nregions
        af = None
        if allele_freq is not None:
            af = torch.as_tensor(allele_freq, dtype=torch.float64).to(self.out.device).contiguous()
            if af.numel() != self.batch.nhaps:
                raise ValueError("allele_freq must hold one frequency per haplotype of the batch")
        flags = torch.empty(max(nreg, 1), dtype=torch.int8, device=self.out.device)
        gaps = torch.empty(2 * max(nreg, 1), dtype=torch.float64, device=self.out.device)
        nflag = ctypes.c_int64()
        code = _lib.lib().vd_genotype_guard(ctypes.byref(self.gb), ctypes.c_void_p(gl.data_ptr()),
                                            ctypes.c_void_p(cert.data_ptr()),
                                            ctypes.c_void_p(af.data_ptr()) if af is not None else None,
                                            float(min_certainty),
                                            float(margin_scale), ctypes.c_void_p(flags.data_ptr()),
                                            ctypes.c_void_p(gaps.data_ptr()), ctypes.c_void_p(s.cuda_stream),
                                            ctypes.byref(nflag))
        _lib.check(code, "vd_genotype_guard")
        report = {"backend": "hip_float64", "reason": "near_tied_likelihood_or_posterior",
                  "margin_scale": margin_scale, "prior": "allele_frequency" if af is not None else "uniform",
                  "flags": flags[:nreg], "gaps": gaps[:2 * nreg], "regions": []}
        if nflag.value == 0:
            return gl, report
        ids = np.flatnonzero(flags[:nreg].cpu().numpy())
        g = gaps[:2 * nreg].cpu().numpy().reshape(-1, 2)
        keep = ~(cert.cpu().numpy() < min_certainty)
        rr = self.batch.region_reads
        mqh = mq.detach().cpu().numpy().astype(np.uint8)
        sub = self.batch.subset(ids)
        dsub = DeviceBatch(sub, device=self.out.device)
        dsub.forward(mode=mode, precision="float64", stream=s)
        smq = torch.from_numpy(np.concatenate([mqh[rr[k]:rr[k + 1]] for k in ids])).to(self.out.device)
        sgl = dsub.genotype(smq, min_certainty, prior, stream=s)
        off = gl_offsets(np.diff(self.batch.region_haps))
        soff = gl_offsets(np.diff(sub.region_haps))
        dst = np.concatenate([np.arange(off[k], off[k + 1]) for k in ids])
        src = np.concatenate([np.arange(soff[i], soff[i + 1]) for i in range(len(ids))])
        gl[torch.from_numpy(dst).to(gl.device)] = sgl[torch.from_numpy(src).to(gl.device)]
        for k in ids:
            mol = int(np.sum(keep[rr[k]:rr[k + 1]]))
            report["regions"].append({"region": int(k), "observed_gaps": [float(g[k, 0]), float(g[k, 1])],
                                      "log_likelihood_margin": margin_scale * max(1, mol)})
        return gl, report

    def result(self) -> np.ndarray:
        return self.out[:self.npairs].cpu().numpy()
