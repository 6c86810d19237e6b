"""CPU FP64 oracle for the PairHMM (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker or the timed CPU
baseline -- never as the thing measured for the GPU, and never as a fallback
for the product path.

* :func:`lib` -- ``liboracle.so``: the plain-C restatement in ``oracle.c``
  (batch_hmm.cpp, trie_vector.cpp, wide.cpp, core.cpp, anchored.cpp of
  caller/native/caller).
* :func:`ref` -- the reference's own C++ sources compiled by
  ``oracle/Makefile`` into ``oracle/_ref`` (absent unless built).
* :mod:`oracle.brute` -- exact rational sum over alignments for tiny pairs.
* :func:`normalized_rows`, :func:`mixture_general`, :func:`mixture_two`,
  :func:`genotype_regions` -- the diploid genotype likelihood chain
  (``genotype.c``: molecule_rows.py, mixture_general.cpp, mixture_two.cpp,
  triangular.py, whole_read.py of caller/native/caller).
"""
from __future__ import annotations

import ctypes
import functools
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_c_char_pp = ctypes.POINTER(ctypes.c_char_p)
_u8p = ctypes.POINTER(ctypes.c_ubyte)
_u8pp = ctypes.POINTER(_u8p)
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


@functools.lru_cache(maxsize=1)
def lib():
    path = HERE / "liboracle.so"
    if not path.exists():
        build()
    L = ctypes.CDLL(str(path))
    L.vdo_error_table.argtypes = [_dp]
    L.vdo_error_table.restype = None
    L.vdo_dense.argtypes = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                            ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int]
    L.vdo_dense.restype = ctypes.c_double
    L.vdo_pairhmm_pairs.argtypes = [_c_char_pp, _u8pp, _ip, _c_char_pp, _ip, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.vdo_pairhmm_pairs.restype = ctypes.c_int
    L.vdo_trie_many.argtypes = [_c_char_pp, _u8pp, _ip, ctypes.c_int, _c_char_pp, _ip, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp,
                                _dp, _u64p]
    L.vdo_trie_many.restype = ctypes.c_int
    L.vdo_regions_cpu.argtypes = [ctypes.c_int, _i64p, _i64p, _i64p, _i64p, _i64p,
                                  ctypes.c_char_p, _u8p, ctypes.c_char_p, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int, _dp, _u64p]
    L.vdo_regions_cpu.restype = ctypes.c_int
    L.vdo_normalized_rows.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _u8p, _dp, _dp]
    L.vdo_normalized_rows.restype = None
    L.vdo_mixture_general.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, _dp]
    L.vdo_mixture_general.restype = None
    L.vdo_mixture_two.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp]
    L.vdo_mixture_two.restype = None
    L.vdo_genotype_regions.argtypes = [ctypes.c_int64, _i64p, ctypes.POINTER(ctypes.c_int32), _i64p, _dp,
                                       _u8p, ctypes.c_double, ctypes.c_double, _i64p, ctypes.c_int, _dp]
    L.vdo_genotype_regions.restype = ctypes.c_int64
    L.vdo_guard.argtypes = [ctypes.c_int64, _i64p, ctypes.POINTER(ctypes.c_int32), _i64p, _i64p, _dp, _dp, _dp,
                            ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int8), _dp]
    L.vdo_guard.restype = ctypes.c_int64
    return L


def error_table() -> np.ndarray:
    out = np.empty(256, np.float64)
    lib().vdo_error_table(out.ctypes.data_as(_dp))
    return out


def _enc(s):
    return s.encode("latin-1") if isinstance(s, str) else bytes(s)


def _qual(q):
    return bytes(q) if isinstance(q, (bytes, bytearray)) else np.asarray(q, dtype=np.uint8).tobytes()


def _err(err):
    e = error_table() if err is None else np.asarray(err, np.float64)
    if e.shape == (64,):  # 64-bin table indexed by min(q, 63)
        e = e[np.minimum(np.arange(256), 63)]
    return np.ascontiguousarray(e, dtype=np.float64)


def dense(read, qual, hap, mode=0, gap=0.001, ext=0.1, err=None, strict=False) -> float:
    """One pair of the dense recurrence (batch_hmm.cpp:31-70; strict: core/wide/anchored)."""
    r, q, h = _enc(read), _qual(qual), _enc(hap)
    e = _err(err)
    qb = (ctypes.c_ubyte * max(1, len(q))).from_buffer_copy(q if q else b"\0")
    return lib().vdo_dense(r, qb, len(r), h, len(h), gap, ext, e.ctypes.data_as(_dp), mode, int(strict))


def pairs(jobs, mode=0, gap=0.001, ext=0.1, err=None, strict=False) -> np.ndarray:
    """Dense recurrence over a list of (read, qual, hap)."""
    return np.array([dense(s, q, h, mode, gap, ext, err, strict) for s, q, h in jobs], np.float64)


def _ptrs(seqs):
    keep = [_enc(s) for s in seqs]
    arr = (ctypes.c_char_p * max(1, len(keep)))(*keep)
    return keep, arr


def trie_many(reads, quals, haps, mode=0, lanes=0, gap=0.001, ext=0.1, err=None):
    """trie_vector.cpp:284 restated; returns (out[nr, nh], cells)."""
    rk, rp = _ptrs(reads)
    qk, qp = _ptrs([_qual(q) for q in quals])
    hk, hp = _ptrs(haps)
    rn = np.array([len(x) for x in rk] or [0], np.int32)
    hn = np.array([len(x) for x in hk] or [0], np.int32)
    out = np.empty((len(rk), len(hk)), np.float64)
    cells = ctypes.c_uint64()
    e = _err(err)
    code = lib().vdo_trie_many(rp, ctypes.cast(qp, _u8pp), rn.ctypes.data_as(_ip), len(rk), hp,
                               hn.ctypes.data_as(_ip), len(hk), mode, lanes, gap, ext,
                               e.ctypes.data_as(_dp), out.ctypes.data_as(_dp), ctypes.byref(cells))
    if code:
        raise RuntimeError("oracle trie_many failed: %d" % code)
    return out, int(cells.value)


def regions_cpu(batch, mode=0, err=None, threads=None):
    """Reference CPU chain (trie + dense fallback) over a RegionBatch on host threads.

    Returns (ln likelihoods in the batch's output layout, number of dense-fallback pairs)."""
    threads = threads or len(os.sched_getaffinity(0))
    out = np.empty(batch.npairs, np.float64)
    e = _err(err)
    dense_pairs = ctypes.c_uint64()
    oo = batch.out_off()
    p = lambda a, t: a.ctypes.data_as(t)
    code = lib().vdo_regions_cpu(
        batch.nregions, p(batch.region_reads, _i64p), p(batch.region_haps, _i64p), p(oo, _i64p),
        p(batch.read_off, _i64p), p(batch.hap_off, _i64p),
        batch.read_bytes.ctypes.data_as(ctypes.c_char_p), p(batch.qual_bytes, _u8p),
        batch.hap_bytes.ctypes.data_as(ctypes.c_char_p), mode, batch.gap, batch.ext,
        p(e, _dp), int(threads), p(out, _dp), ctypes.byref(dense_pairs))
    if code:
        raise RuntimeError("oracle regions_cpu failed: %d" % code)
    return out, int(dense_pairs.value)


# ---------------------------------------------------------------------------
# diploid genotype likelihood (genotype.c)

def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def normalized_rows(ll, mq):
    """molecule_rows.py:6-14 -> (values, certainty)."""
    x = _f64(ll)
    rows, n = x.shape
    q = np.ascontiguousarray(mq, np.uint8)
    values = np.empty_like(x)
    cert = np.empty(rows, np.float64)
    if rows and n:
        lib().vdo_normalized_rows(rows, n, x.ctypes.data_as(_dp), q.ctypes.data_as(_u8p),
                                  values.ctypes.data_as(_dp), cert.ctypes.data_as(_dp))
    return values, cert


def mixture_general(lh, prior=None, triangular=True):
    """mixture_general.cpp:9-40 (prior None -> -1, as mixture_cpu.py:27)."""
    x = _f64(lh)
    rows, n = x.shape
    out = np.empty(n * (n + 1) // 2 if triangular else n * n)
    lib().vdo_mixture_general(rows, n, x.ctypes.data_as(_dp), -1.0 if prior is None else float(prior),
                              int(bool(triangular)), out.ctypes.data_as(_dp))
    return out


def mixture_two(lh, prior):
    """mixture_two.cpp:9-29 -> 2 x 2 ordered genotypes."""
    x = _f64(lh)
    out = np.empty(4)
    lib().vdo_mixture_two(len(x), x.ctypes.data_as(_dp), float(prior), out.ctypes.data_as(_dp))
    return out.reshape(2, 2)


def diploid_likelihood(lh, artifact_prior=None):
    """triangular.py:6-11."""
    x = _f64(lh)
    if x.shape[1] == 2 and artifact_prior is not None:
        return mixture_two(x, artifact_prior).ravel()[[0, 1, 3]]
    return mixture_general(x, artifact_prior, True)


def genotype_regions(row_off, ncol, mat_off, ll, mq, gl_off, min_certainty=float(np.log(3)),
                     prior=0.001, threads=None):
    """whole_read.py:30-49 single-end chain over a region batch -> (gl, invalid regions)."""
    row_off = np.ascontiguousarray(row_off, np.int64)
    ncol = np.ascontiguousarray(ncol, np.int32)
    mat_off = np.ascontiguousarray(mat_off, np.int64)
    gl_off = np.ascontiguousarray(gl_off, np.int64)
    ll = _f64(ll)
    mq = np.ascontiguousarray(mq, np.uint8)
    nreg = len(ncol)
    total = int(gl_off[-1]) if len(gl_off) > nreg else int(
        gl_off[-1] + ncol[-1] * (ncol[-1] + 1) // 2) if nreg else 0
    gl = np.full(max(total, 1), np.nan)
    if threads is None:
        threads = os.cpu_count() or 1
    bad = lib().vdo_genotype_regions(nreg, row_off.ctypes.data_as(_i64p),
                                     ncol.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     mat_off.ctypes.data_as(_i64p), ll.ctypes.data_as(_dp),
                                     mq.ctypes.data_as(_u8p), float(min_certainty), float(prior),
                                     gl_off.ctypes.data_as(_i64p), int(threads), gl.ctypes.data_as(_dp))
    return gl[:total], int(bad)


def guard(row_off, ncol, gl_off, gl, cert, af=None, min_certainty=float(np.log(3)), margin_scale=1e-4):
    """partial_anchor.py:119-127 near-tie guard (prior from allele frequencies af,
    one per haplotype column, partial_anchor.py:56-57; None: the uniform prior of
    whole_read.py:51) -> (flags int8 [nregions], gaps float64 [nregions, 2])."""
    row_off = np.ascontiguousarray(row_off, np.int64)
    ncol = np.ascontiguousarray(ncol, np.int32)
    gl_off = np.ascontiguousarray(gl_off, np.int64)
    col_off = np.concatenate([[0], np.cumsum(ncol)]).astype(np.int64)
    gl, cert = _f64(gl), _f64(cert)
    nreg = len(ncol)
    flags = np.zeros(max(nreg, 1), np.int8)
    gaps = np.zeros(2 * max(nreg, 1), np.float64)
    a = _f64(af) if af is not None else None
    lib().vdo_guard(nreg, row_off.ctypes.data_as(_i64p), ncol.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                    col_off.ctypes.data_as(_i64p), gl_off.ctypes.data_as(_i64p), gl.ctypes.data_as(_dp),
                    cert.ctypes.data_as(_dp), a.ctypes.data_as(_dp) if a is not None else None,
                    float(min_certainty), float(margin_scale),
                    flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), gaps.ctypes.data_as(_dp))
    return flags[:nreg], gaps[:2 * nreg].reshape(-1, 2)


# ---------------------------------------------------------------------------
# the reference's own sources compiled into oracle/_ref (never shipped)

class Ref:
    """ctypes bindings of oracle/_ref/lib*.so with the reference's argtypes
    (cpu_batch.py:14-17, trie_vector.py:11-13, whole_pairhmm.py:38-46,
    base/caller.py:40-42, indel_refine.py:56-57)."""

    def __init__(self, root: Path):
        self.batch = ctypes.CDLL(str(root / "libbatch_hmm.so"))
        self.batch.pairhmm_pairs.argtypes = [_c_char_pp, _u8pp, _ip, _c_char_pp, _ip, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, _dp]
        self.batch.pairhmm_pairs.restype = ctypes.c_int
        self.trie = ctypes.CDLL(str(root / "libtrie_vector.so"))
        self.trie.trie_many.argtypes = [_c_char_pp, _u8pp, _ip, ctypes.c_int, _c_char_pp, _ip,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _u64p]
        self.trie.trie_many.restype = ctypes.c_int
        args7 = [ctypes.c_char_p, _u8p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                 ctypes.c_double, ctypes.c_double]
        self.wide = ctypes.CDLL(str(root / "libwide.so"))
        self.wide.read_haplotype.argtypes = args7
        self.wide.read_haplotype.restype = ctypes.c_double
        batch_args = [ctypes.c_char_p, _u8p, ctypes.c_int, _c_char_pp, _ip, ctypes.c_int,
                      ctypes.c_double, ctypes.c_double, _dp]
        self.wide.read_haplotype_batch_mode.argtypes = batch_args + [ctypes.c_int]
        self.wide.read_haplotype_batch_mode.restype = None
        self.core = ctypes.CDLL(str(root / "libcore.so"))
        self.core.read_haplotype.argtypes = args7
        self.core.read_haplotype.restype = ctypes.c_double
        self.core.read_haplotype_batch.argtypes = batch_args
        self.core.read_haplotype_batch.restype = None
        self.anchored = ctypes.CDLL(str(root / "libanchored.so"))
        self.anchored.anchored_pairhmm.argtypes = args7
        self.anchored.anchored_pairhmm.restype = ctypes.c_double
        self.mix_two = self.mix_general = None
        if (root / "libmixture_two.so").exists():   # mixture_cpu.py:10,20
            self.mix_two = ctypes.CDLL(str(root / "libmixture_two.so"))
            self.mix_two.mixture_two.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp]
            self.mix_two.mixture_two.restype = None
        if (root / "libmixture_general.so").exists():
            self.mix_general = ctypes.CDLL(str(root / "libmixture_general.so"))
            self.mix_general.mixture_general.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                                                         ctypes.c_int, _dp]
            self.mix_general.mixture_general.restype = None

    # convenience wrappers ------------------------------------------------
    def pairhmm_pairs(self, jobs, mode, lanes=0):
        rk, rp = _ptrs([s for s, q, h in jobs])
        qk, qp = _ptrs([_qual(q) for s, q, h in jobs])
        hk, hp = _ptrs([h for s, q, h in jobs])
        rn = np.array([len(x) for x in rk] or [0], np.int32)
        hn = np.array([len(x) for x in hk] or [0], np.int32)
        out = np.empty(len(jobs), np.float64)
        code = self.batch.pairhmm_pairs(rp, ctypes.cast(qp, _u8pp), rn.ctypes.data_as(_ip), hp,
                                        hn.ctypes.data_as(_ip), len(jobs), mode, lanes,
                                        out.ctypes.data_as(_dp))
        if code:
            raise RuntimeError("CPU PairHMM failed: %d" % code)
        return out

    def trie_many(self, reads, quals, haps, mode, lanes=0):
        rk, rp = _ptrs(reads)
        qk, qp = _ptrs([_qual(q) for q in quals])
        hk, hp = _ptrs(haps)
        rn = np.array([len(x) for x in rk] or [0], np.int32)
        hn = np.array([len(x) for x in hk] or [0], np.int32)
        out = np.empty((len(rk), len(hk)), np.float64)
        cells = ctypes.c_uint64()
        code = self.trie.trie_many(rp, ctypes.cast(qp, _u8pp), rn.ctypes.data_as(_ip), len(rk), hp,
                                   hn.ctypes.data_as(_ip), len(hk), mode, lanes,
                                   out.ctypes.data_as(_dp), ctypes.byref(cells))
        if code:
            raise RuntimeError("Vector prefix PairHMM failed: %d" % code)
        return out, int(cells.value)

    def mixture_general(self, lh, prior=None, triangular=True):
        x = _f64(lh)
        rows, n = x.shape
        out = np.empty(n * (n + 1) // 2 if triangular else n * n)
        self.mix_general.mixture_general(rows, n, x.ctypes.data_as(_dp), -1.0 if prior is None else float(prior),
                                         int(bool(triangular)), out.ctypes.data_as(_dp))
        return out

    def mixture_two(self, lh, prior):
        x = _f64(lh)
        out = np.empty(4)
        self.mix_two.mixture_two(len(x), x.ctypes.data_as(_dp), float(prior), out.ctypes.data_as(_dp))
        return out.reshape(2, 2)

    def scalar(self, fn, jobs, gap=0.001, ext=0.1):
        f = {"wide": self.wide.read_haplotype, "core": self.core.read_haplotype,
             "anchored": self.anchored.anchored_pairhmm}[fn]
        res = []
        for s, q, h in jobs:
            qb = _qual(q)
            qa = (ctypes.c_ubyte * max(1, len(qb))).from_buffer_copy(qb if qb else b"\0")
            res.append(f(_enc(s), qa, len(_enc(s)), _enc(h), len(_enc(h)), gap, ext))
        return np.array(res, np.float64)


@functools.lru_cache(maxsize=1)
def ref():
    """The reference compiled from source, or None when oracle/_ref is absent."""
    root = HERE / "_ref"
    if not (root / "libtrie_vector.so").exists():
        return None
    return Ref(root)
