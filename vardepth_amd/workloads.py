"""Region batches (the engine's input layout) and seeded synthetic workloads.

A :class:`RegionBatch` is the CSR form of many active regions, each with its
own reads, base qualities and candidate haplotypes -- the same data that the
reference's region call ``pair_scores_many(lib, sequences, qualities, haps)``
(caller/native/caller/base/caller.py:151-160) receives one region at a time.
Output for region ``k`` is the row-major ``R_k x H_k`` natural-log likelihood
block at ``out_off[k]``.

The generators restate the benchmark workloads as concrete seeded inputs
(numpy PCG64):

* :func:`single_region` -- one active region, 64 reads x 8 haplotypes of
  200-400 bp, 150 bp reads, Phred U{10..40}, gap open Q45, continuation Q10.
* :func:`uneven_regions` -- 4096 regions, 2-32 haplotypes of 150-600 bp,
  low/high-depth read counts, lognormal read lengths clipped to 50-250 bp,
  bimodal Phred (80% in 30-40, 20% in 2-20).
* :func:`precision_stress` -- 1024 regions of 250 bp reads against 1-2 kb
  indel-dense haplotypes at Phred 2-15.
* :func:`hmm_sampled_regions` -- reads sampled from the PairHMM itself (so
  gap parameters are identifiable), with the true haplotype of every read.
* :func:`chromosome_sample` -- a seeded random chromosome with heterozygous
  SNVs and indels on two haplotypes, 150 bp reads at a chosen depth, and
  400 bp active regions whose candidate haplotypes are every combination of
  the variant and candidate sites inside them (1..64 haplotypes).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BASES = np.frombuffer(b"ACGT", dtype=np.uint8)
GAP_Q45 = 10.0 ** -4.5      # "indel quality Q45 open"
EXT_Q10 = 10.0 ** -1.0      # "Q10 continuation"


@dataclasses.dataclass
class RegionBatch:
    """CSR batch of active regions (all offsets int64, all bases uint8)."""

    read_bytes: np.ndarray
    qual_bytes: np.ndarray
    read_off: np.ndarray
    hap_bytes: np.ndarray
    hap_off: np.ndarray
    region_reads: np.ndarray
    region_haps: np.ndarray
    gap: float = 0.001
    ext: float = 0.1
    read_true_hap: np.ndarray | None = None   # global hap index per read (or None)

    def __post_init__(self):
        self.read_bytes = np.ascontiguousarray(self.read_bytes, dtype=np.uint8)
        self.qual_bytes = np.ascontiguousarray(self.qual_bytes, dtype=np.uint8)
        self.read_off = np.ascontiguousarray(self.read_off, dtype=np.int64)
        self.hap_bytes = np.ascontiguousarray(self.hap_bytes, dtype=np.uint8)
        self.hap_off = np.ascontiguousarray(self.hap_off, dtype=np.int64)
        self.region_reads = np.ascontiguousarray(self.region_reads, dtype=np.int64)
        self.region_haps = np.ascontiguousarray(self.region_haps, dtype=np.int64)
        if len(self.read_bytes) != len(self.qual_bytes):
            raise ValueError("read and quality byte arrays differ in length")
        if self.read_off[-1] != len(self.read_bytes) or self.hap_off[-1] != len(self.hap_bytes):
            raise ValueError("offsets do not cover the byte arrays")
        if len(self.region_reads) != len(self.region_haps):
            raise ValueError("region read/haplotype ranges differ in count")

    # -- shape helpers -------------------------------------------------
    @property
    def nregions(self) -> int:
        return len(self.region_reads) - 1

    @property
    def nreads(self) -> int:
        return len(self.read_off) - 1

    @property
    def nhaps(self) -> int:
        return len(self.hap_off) - 1

    def read_lengths(self) -> np.ndarray:
        return np.diff(self.read_off)

    def hap_lengths(self) -> np.ndarray:
        return np.diff(self.hap_off)

    def region_counts(self):
        return np.diff(self.region_reads), np.diff(self.region_haps)

    def out_off(self) -> np.ndarray:
        r, h = self.region_counts()
        out = np.zeros(self.nregions + 1, dtype=np.int64)
        np.cumsum(r * h, out=out[1:])
        return out

    @property
    def npairs(self) -> int:
        return int(self.out_off()[-1])

    def cells(self) -> int:
        """Sum over pairs of n*m (the reference's dp_cells, accelerator.py:155)."""
        a = self.read_off[self.region_reads[1:]] - self.read_off[self.region_reads[:-1]]
        b = self.hap_off[self.region_haps[1:]] - self.hap_off[self.region_haps[:-1]]
        return int(np.sum(a * b))

    def region(self, k):
        """(reads, quals, haps) of region k as Python str/bytes lists."""
        r0, r1 = self.region_reads[k], self.region_reads[k + 1]
        h0, h1 = self.region_haps[k], self.region_haps[k + 1]
        reads = [self.read_bytes[self.read_off[i]:self.read_off[i + 1]].tobytes().decode("latin-1")
                 for i in range(r0, r1)]
        quals = [self.qual_bytes[self.read_off[i]:self.read_off[i + 1]].tobytes() for i in range(r0, r1)]
        haps = [self.hap_bytes[self.hap_off[i]:self.hap_off[i + 1]].tobytes().decode("latin-1")
                for i in range(h0, h1)]
        return reads, quals, haps

    def subset(self, regions) -> "RegionBatch":
        """A new batch holding the given regions (in the given order)."""
        return from_regions([self.region(k) for k in regions], gap=self.gap, ext=self.ext)


def from_regions(regions, gap=0.001, ext=0.1) -> RegionBatch:
    """Build a batch from a list of (reads, quals, haps) Python lists."""
    rb, qb, ro, hb, ho, rr, rh = [], [], [0], [], [0], [0], [0]
    for reads, quals, haps in regions:
        if len(reads) != len(quals):
            raise ValueError("Sequence/quality batch lengths differ")
        for s, q in zip(reads, quals):
            s = s.encode("latin-1") if isinstance(s, str) else bytes(s)
            q = bytes(q) if isinstance(q, (bytes, bytearray)) else np.asarray(q, dtype=np.uint8).tobytes()
            if len(s) != len(q):
                raise ValueError("Read and quality lengths differ")
            rb.append(s)
            qb.append(q)
            ro.append(ro[-1] + len(s))
        for h in haps:
            h = h.encode("latin-1") if isinstance(h, str) else bytes(h)
            hb.append(h)
            ho.append(ho[-1] + len(h))
        rr.append(rr[-1] + len(reads))
        rh.append(rh[-1] + len(haps))
    return RegionBatch(
        read_bytes=np.frombuffer(b"".join(rb), dtype=np.uint8).copy(),
        qual_bytes=np.frombuffer(b"".join(qb), dtype=np.uint8).copy(),
        read_off=np.array(ro, dtype=np.int64),
        hap_bytes=np.frombuffer(b"".join(hb), dtype=np.uint8).copy(),
        hap_off=np.array(ho, dtype=np.int64),
        region_reads=np.array(rr, dtype=np.int64),
        region_haps=np.array(rh, dtype=np.int64),
        gap=gap, ext=ext)


# ---------------------------------------------------------------------------
# generators

def _mutate(rng, base: np.ndarray, snvs: int, indels: int, max_indel: int = 10) -> np.ndarray:
    """Copy of ``base`` with ``snvs`` substitutions and ``indels`` indels of 1..max_indel bp."""
    h = base.copy()
    for _ in range(snvs):
        p = int(rng.integers(0, len(h)))
        h[p] = BASES[(np.searchsorted(BASES, h[p]) + int(rng.integers(1, 4))) % 4]
    for _ in range(indels):
        size = int(rng.integers(1, max_indel + 1))
        p = int(rng.integers(1, max(2, len(h) - size)))
        if rng.random() < 0.5:
            h = np.concatenate([h[:p], BASES[rng.integers(0, 4, size)], h[p:]])
        else:
            h = np.concatenate([h[:p], h[p + size:]])
    return h


def _substitute(rng, seq: np.ndarray, qual: np.ndarray) -> np.ndarray:
    """Substitute each base with probability 10^(-q/10) to one of the other three."""
    err = np.power(10.0, -qual.astype(np.float64) / 10.0)
    hit = rng.random(len(seq)) < err
    if hit.any():
        code = np.searchsorted(BASES, seq[hit])
        seq = seq.copy()
        seq[hit] = BASES[(code + rng.integers(1, 4, int(hit.sum()))) % 4]
    return seq


class _Reads:
    """Reads of one region as concatenated bases / qualities plus lengths."""

    def __init__(self, seq, qual, lengths):
        self.seq, self.qual, self.lengths = seq, qual, np.asarray(lengths, dtype=np.int64)

    def __len__(self):
        return len(self.lengths)


class _Builder:
    def __init__(self):
        self.rseq, self.rqual, self.rlen, self.haps = [], [], [], []
        self.rcount, self.hcount, self.truth = [], [], []
        self.nhaps = 0

    def add(self, haps, reads, quals, truth=None):
        if not isinstance(reads, _Reads):
            reads = _Reads(np.concatenate(reads) if reads else np.zeros(0, np.uint8),
                           np.concatenate(quals) if quals else np.zeros(0, np.uint8),
                           [len(r) for r in reads])
        self.rseq.append(reads.seq)
        self.rqual.append(reads.qual)
        self.rlen.append(reads.lengths)
        self.haps.extend(haps)
        self.rcount.append(len(reads))
        self.hcount.append(len(haps))
        if truth is not None:
            self.truth.extend(int(t) + self.nhaps for t in truth)
        self.nhaps += len(haps)

    def batch(self, gap, ext) -> RegionBatch:
        def cat(arrs):
            if not arrs:
                return np.zeros(0, np.uint8), np.zeros(1, np.int64)
            off = np.zeros(len(arrs) + 1, dtype=np.int64)
            np.cumsum([len(a) for a in arrs], out=off[1:])
            return np.concatenate(arrs).astype(np.uint8), off
        rb = np.concatenate(self.rseq).astype(np.uint8) if self.rseq else np.zeros(0, np.uint8)
        qb = np.concatenate(self.rqual).astype(np.uint8) if self.rqual else np.zeros(0, np.uint8)
        lens = np.concatenate(self.rlen) if self.rlen else np.zeros(0, np.int64)
        ro = np.zeros(len(lens) + 1, dtype=np.int64)
        np.cumsum(lens, out=ro[1:])
        hb, ho = cat(self.haps)
        rr = np.zeros(len(self.rcount) + 1, dtype=np.int64)
        rh = np.zeros(len(self.hcount) + 1, dtype=np.int64)
        np.cumsum(self.rcount, out=rr[1:])
        np.cumsum(self.hcount, out=rh[1:])
        truth = np.array(self.truth, dtype=np.int64) if self.truth else None
        return RegionBatch(rb, qb, ro, hb, ho, rr, rh, gap=gap, ext=ext, read_true_hap=truth)


def _region_haps(rng, length, count, snv_max=5, indel_max=2, max_indel=10):
    base = BASES[rng.integers(0, 4, length)]
    return [_mutate(rng, base, int(rng.integers(0, snv_max + 1)), int(rng.integers(0, indel_max + 1)),
                    max_indel) for _ in range(count)]


def _sample_reads(rng, haps, lengths, qual_fn):
    """Reads of the given lengths sampled from random haplotypes at random offsets,
    with qualities qual_fn(total) and substitution errors (vectorised per region)."""
    lengths = np.asarray(lengths, dtype=np.int64)
    if len(lengths) == 0:
        return [], [], []
    hcat = np.concatenate(haps)
    hoff = np.zeros(len(haps) + 1, dtype=np.int64)
    np.cumsum([len(h) for h in haps], out=hoff[1:])
    hl = np.diff(hoff)
    src = rng.integers(0, len(haps), len(lengths))
    n = np.minimum(lengths, hl[src])
    o = (rng.random(len(n)) * (hl[src] - n + 1)).astype(np.int64)
    starts = hoff[src] + o
    roff = np.zeros(len(n) + 1, dtype=np.int64)
    np.cumsum(n, out=roff[1:])
    idx = np.repeat(starts - roff[:-1], n) + np.arange(roff[-1])
    q = qual_fn(int(roff[-1]))
    seq = _substitute(rng, hcat[idx], q)
    return _Reads(seq, q, n), None, src.tolist()


def single_region(seed: int = 41001, reads: int = 64, haps: int = 8, read_len: int = 150,
                  gap: float = GAP_Q45, ext: float = EXT_Q10) -> RegionBatch:
    """One active region: 64 reads x 8 haplotypes of 200-400 bp."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = _Builder()
    length = int(rng.integers(200, 401))
    hs = _region_haps(rng, length, haps)
    rs, qs, truth = _sample_reads(rng, hs, [read_len] * reads,
                                  lambda n: rng.integers(10, 41, n).astype(np.uint8))
    b.add(hs, rs, qs, truth)
    return b.batch(gap, ext)


def _bimodal_quals(rng):
    def fn(n):
        hi = rng.random(n) < 0.8
        q = np.where(hi, rng.integers(30, 41, n), rng.integers(2, 21, n))
        return q.astype(np.uint8)
    return fn


def uneven_regions(seed: int = 41002, regions: int = 4096, gap: float = 0.001,
                   ext: float = 0.1) -> RegionBatch:
    """4096 uneven active regions (2-32 haplotypes of 150-600 bp; 2-30 or 200-2000 reads)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = _Builder()
    qfn = _bimodal_quals(rng)
    for _ in range(regions):
        h = int(rng.integers(2, 33))
        length = int(rng.integers(150, 601))
        hs = _region_haps(rng, length, h)
        r = int(rng.integers(2, 31)) if rng.random() < 0.85 else int(rng.integers(200, 2001))
        lens = np.clip(np.rint(rng.lognormal(math.log(150.0), 0.35, r)), 50, 250).astype(int)
        rs, qs, truth = _sample_reads(rng, hs, lens, qfn)
        b.add(hs, rs, qs, truth)
    return b.batch(gap, ext)


def precision_stress(seed: int = 41003, regions: int = 1024, reads: int = 16,
                     gap: float = GAP_Q45, ext: float = EXT_Q10) -> RegionBatch:
    """250 bp low-quality reads against 1-2 kb haplotypes with one indel per 50 bp."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = _Builder()
    for _ in range(regions):
        h = int(rng.integers(2, 5))
        length = int(rng.integers(1000, 2001))
        base = BASES[rng.integers(0, 4, length)]
        hs = [_mutate(rng, base, int(rng.integers(0, 6)), max(1, length // 50), 10) for _ in range(h)]
        rs, qs, truth = _sample_reads(rng, hs, [250] * reads,
                                      lambda n: rng.integers(2, 16, n).astype(np.uint8))
        b.add(hs, rs, qs, truth)
    return b.batch(gap, ext)


def _hmm_reads(rng, hcat: np.ndarray, hoff: np.ndarray, src: np.ndarray, n: int, qual: np.ndarray, gap: float,
               ext: float) -> np.ndarray:
    """Sample reads of n bases from the PairHMM, all reads at once: each starts
    uniformly in its haplotype ``src[r]``; M->I / M->D with probability gap,
    I->I / D->D with probability ext; match emission 1-e else a uniform other
    base, insertions uniform over ACGT; a deletion running off the haplotype end
    returns to M on the last base."""
    R = len(src)
    hl = hoff[src + 1] - hoff[src]
    j = (rng.random(R) * np.maximum(1, hl - n + 1)).astype(np.int64)
    state = np.zeros(R, np.int8)  # 0 = M, 1 = I, 2 = D
    i = np.zeros(R, np.int64)
    out = np.zeros((R, n), np.uint8)
    err = np.power(10.0, -qual.astype(np.float64) / 10.0)
    rows = np.arange(R)
    for _ in range(8 * n):
        act = i < n
        if not act.any():
            break
        u = rng.random(R)
        nxt = np.where(state == 0, np.where(u < gap, 1, np.where(u < 2 * gap, 2, 0)), np.where(u < ext, state, 0))
        state = np.where(act, nxt, state).astype(np.int8)
        dele = act & (state == 2)
        j = np.where(dele, j + 1, j)
        off_end = dele & (j >= hl)
        state = np.where(off_end, 0, state).astype(np.int8)
        ins = act & (state == 1)
        mat = act & (state == 0) & ~dele
        ii = np.minimum(i, n - 1)
        rnd = BASES[rng.integers(0, 4, R)]
        jj = np.clip(j, 0, hl - 1)
        base = hcat[hoff[src] + jj]
        sub = rng.random(R) < err[rows, ii]
        shifted = BASES[(np.searchsorted(BASES, base) + rng.integers(1, 4, R)) % 4]
        emit = np.where(ins, rnd, np.where(sub, shifted, base))
        emitting = ins | mat
        out[rows[emitting], ii[emitting]] = emit[emitting]
        i = np.where(emitting, i + 1, i)
        j = np.where(mat, jj + 1, j)
    return out


def hmm_sampled_regions(seed: int = 41004, regions: int = 1024, reads: int = 100, haps: int = 20,
                        read_len: int = 150, gap: float = GAP_Q45, ext: float = EXT_Q10,
                        sample_gap: float = 0.01, sample_ext: float = 0.3) -> RegionBatch:
    """Regions whose reads are drawn from the PairHMM itself (gap sample_gap,
    continuation sample_ext, reference error table), true haplotype per read
    recorded; haplotypes as in single_region (200-400 bp windows with SNVs and
    indels); qualities uniform 10..40."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hs_all, rcount = [], []
    for _ in range(regions):
        hs_all.append(_region_haps(rng, int(rng.integers(200, 401)), haps))
        rcount.append(reads)
    hlen = np.array([len(h) for hs in hs_all for h in hs], np.int64)
    hoff = np.zeros(len(hlen) + 1, np.int64)
    np.cumsum(hlen, out=hoff[1:])
    hcat = np.concatenate([h for hs in hs_all for h in hs])
    nreads = regions * reads
    local = rng.integers(0, haps, nreads)
    src = np.repeat(np.arange(regions) * haps, reads) + local
    qual = rng.integers(10, 41, (nreads, read_len)).astype(np.uint8)
    seq = _hmm_reads(rng, hcat, hoff, src, read_len, qual, sample_gap, sample_ext)
    b = _Builder()
    for k in range(regions):
        r0 = k * reads
        rs = _Reads(seq[r0:r0 + reads].reshape(-1), qual[r0:r0 + reads].reshape(-1), [read_len] * reads)
        b.add(hs_all[k], rs, None, list(local[r0:r0 + reads]))
    return b.batch(gap, ext)


def true_pairs(batch: RegionBatch) -> RegionBatch:
    """The batch restricted to each read's true haplotype: one region per
    (region, haplotype) with the reads sampled from it."""
    if batch.read_true_hap is None:
        raise ValueError("batch has no true haplotype per read")
    regions = []
    for k in range(batch.nregions):
        reads, quals, haps = batch.region(k)
        r0, h0 = batch.region_reads[k], batch.region_haps[k]
        truth = batch.read_true_hap[r0:batch.region_reads[k + 1]] - h0
        for h in range(len(haps)):
            sel = np.flatnonzero(truth == h)
            if len(sel):
                regions.append(([reads[i] for i in sel], [quals[i] for i in sel], [haps[h]]))
    return from_regions(regions, batch.gap, batch.ext)


# ---------------------------------------------------------------------------
# chromosome-scale diploid sample

@dataclasses.dataclass
class ChromosomeSample:
    """A region batch with what the genotype chain needs besides it."""

    batch: RegionBatch
    mq: np.ndarray           # uint8 mapping quality of every batch read
    true_pair: np.ndarray    # per region: triangular index of the true genotype (b = 0..n-1, a = 0..b)
    region_start: np.ndarray  # per region: reference start of the window
    info: dict


def _apply_sites(ref: np.ndarray, lo: int, hi: int, pos, rlen, alt, pieces_out=None):
    """ref[lo:hi] with the (sorted, non-overlapping) sites applied: site i
    replaces ref[pos[i] : pos[i] + rlen[i]] by the bytes alt[i]."""
    out, cur = [], lo
    for p, r, a in zip(pos, rlen, alt):
        out.append(ref[cur:p])
        if pieces_out is not None:
            pieces_out.append((cur, p - cur, True))
        out.append(a)
        if pieces_out is not None:
            pieces_out.append((p, len(a), False))
        cur = p + r
    out.append(ref[cur:hi])
    if pieces_out is not None:
        pieces_out.append((cur, hi - cur, True))
    return np.concatenate(out) if len(out) > 1 else ref[lo:hi].copy()


def chromosome_sample(seed: int = 41005, mb: float = 60.0, depth: float = 30.0, read_len: int = 150,
                      window: int = 400, padding: int = 200, snv_rate: float = 1 / 1500,
                      indel_rate: float = 1 / 10000, max_indel: int = 10, false_sites: float = 1.0,
                      max_haps: int = 64, gap: float = GAP_Q45, ext: float = EXT_Q10,
                      progress=None) -> ChromosomeSample:
    """Seeded chromosome-scale diploid sample.

    * reference: ``mb`` million uniform random bases;
    * two haplotypes: heterozygous SNVs at ``snv_rate`` and indels (1..max_indel
      bp, insertion or deletion) at ``indel_rate`` per base, each on one
      haplotype, sites at least 12 bp apart;
    * reads: ``read_len`` bp at ``depth`` (half from each haplotype) at uniform
      starts, Phred 85% U{30..40} / 15% U{10..29} with substitutions drawn
      from them, mapping quality 60 (95%) or U{20..59};
    * regions: consecutive ``window`` bp windows; candidate sites = the true
      sites in the window plus Poisson(``false_sites``) candidate SNVs; the
      haplotypes are ref[start - padding : end + padding] with every subset of
      the sites applied (subset mask order, so haplotype 0 is the reference),
      trimmed to at most ``max_haps`` by dropping candidate sites first;
    * ``progress`` (optional callable) receives a short message per stage;
    * each region holds the reads overlapping its window, in start order
      (whole_read.py:16-27 fetches the overlapping reads; with reads shorter
      than the padding every one lies inside the padded window).
    """
    rng = np.random.default_rng(seed)
    L = int(round(mb * 1e6))
    ref = BASES[rng.integers(0, 4, L)]
    lo_ok, hi_ok = padding + 16, L - padding - 16
    # --- heterozygous sites
    nsnv = rng.poisson(L * snv_rate)
    nind = rng.poisson(L * indel_rate)
    pos = np.concatenate([rng.integers(lo_ok, hi_ok, nsnv), rng.integers(lo_ok, hi_ok, nind)])
    kind = np.concatenate([np.zeros(nsnv, np.int8), np.ones(nind, np.int8)])
    order = np.argsort(pos, kind="stable")
    pos, kind = pos[order], kind[order]
    keep = np.ones(len(pos), bool)
    keep[1:] = np.diff(pos) >= 12
    pos, kind = pos[keep], kind[keep]
    nsite = len(pos)
    phase = rng.integers(0, 2, nsite).astype(np.int8)
    rlen = np.ones(nsite, np.int64)
    alt = []
    dl = rng.integers(1, max_indel + 1, nsite)
    ins = rng.random(nsite) < 0.5
    for i in range(nsite):
        r = ref[pos[i]]
        if kind[i] == 0:
            alt.append(np.array([BASES[(int(np.flatnonzero(BASES == r)[0]) + rng.integers(1, 4)) % 4]], np.uint8))
        elif ins[i]:
            alt.append(np.concatenate([[r], BASES[rng.integers(0, 4, dl[i])]]).astype(np.uint8))
        else:
            rlen[i] = 1 + dl[i]
            alt.append(np.array([r], np.uint8))
    # --- the two haplotypes and their piece maps back to the reference
    haps, maps = [], []
    for h in (0, 1):
        sel = np.flatnonzero(phase == h)
        pieces = []
        seq = _apply_sites(ref, 0, L, pos[sel], rlen[sel], [alt[i] for i in sel], pieces)
        ref_start = np.array([p[0] for p in pieces], np.int64)
        plen = np.array([p[1] for p in pieces], np.int64)
        is_ref = np.array([p[2] for p in pieces], bool)
        hap_start = np.concatenate([[0], np.cumsum(plen)[:-1]])
        haps.append(seq)
        maps.append((hap_start, ref_start, is_ref))
    if progress:
        progress("chromosome sample: %d sites on two haplotypes" % nsite)
    # --- reads
    rs_all, seqs, quals = [], [], []
    nper = int(round(depth * L / (2 * read_len)))
    for h in (0, 1):
        seq = haps[h]
        hs, rstart, isref = maps[h]
        st = np.sort(rng.integers(0, len(seq) - read_len + 1, nper))
        piece = np.searchsorted(hs, st, side="right") - 1
        rs = np.where(isref[piece], rstart[piece] + (st - hs[piece]), rstart[piece])
        rs_all.append(rs)
        win = np.lib.stride_tricks.sliding_window_view(seq, read_len)
        for c0 in range(0, nper, 1 << 19):
            idx = st[c0:c0 + (1 << 19)]
            r = win[idx].copy()
            q = np.where(rng.random(r.shape, dtype=np.float32) < 0.85,
                         rng.integers(30, 41, r.shape, dtype=np.uint8),
                         rng.integers(10, 30, r.shape, dtype=np.uint8))
            err = rng.random(r.shape, dtype=np.float32) < (10.0 ** (-q.astype(np.float32) / 10))
            if err.any():
                code = np.searchsorted(BASES, r[err])
                r[err] = BASES[(code + rng.integers(1, 4, code.shape)) % 4]
            seqs.append(r)
            quals.append(q)
    if progress:
        progress("chromosome sample: %d reads" % (2 * nper))
    rs = np.concatenate(rs_all)
    rseq = np.concatenate(seqs)
    rqual = np.concatenate(quals)
    nreads_total = len(rs)
    read_mq = np.where(rng.random(nreads_total) < 0.95, 60, rng.integers(20, 60, nreads_total)).astype(np.uint8)
    o = np.argsort(rs, kind="stable")
    rs, rseq, rqual, read_mq = rs[o], rseq[o], rqual[o], read_mq[o]
    # --- windows: reads overlapping [w*window, (w+1)*window)
    nwin = (L - 2 * (padding + 16)) // window
    w0 = padding + 16
    wstart = w0 + np.arange(nwin, dtype=np.int64) * window
    first = np.searchsorted(rs, wstart - read_len + 1, side="left")
    last = np.searchsorted(rs, wstart + window, side="left")
    counts = last - first
    region_reads = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ridx = np.concatenate([np.arange(a, b) for a, b in zip(first, last)]) if nwin else np.zeros(0, np.int64)
    read_bytes = rseq[ridx].reshape(-1)
    qual_bytes = rqual[ridx].reshape(-1)
    read_off = np.arange(len(ridx) + 1, dtype=np.int64) * read_len
    mq = read_mq[ridx]
    # --- candidate haplotypes per window
    site_w = np.searchsorted(wstart, pos, side="right") - 1
    hb, hlen, nh, true_pair = [], [], [], []
    nfalse = rng.poisson(false_sites, nwin)
    kmax = int(math.floor(math.log2(max_haps)))
    for w in range(nwin):
        if progress and w and w % 25000 == 0:
            progress("chromosome sample: haplotypes of %d / %d regions" % (w, nwin))
        s0, s1 = int(wstart[w]), int(wstart[w] + window)
        tsel = np.flatnonzero((site_w == w) & (pos + rlen <= s1 + padding))
        cpos = [int(x) for x in pos[tsel]]
        crl = [int(x) for x in rlen[tsel]]
        calt = [alt[i] for i in tsel]
        cph = [int(phase[i]) for i in tsel]
        for _ in range(int(nfalse[w])):      # candidate SNVs away from every other site
            p = int(rng.integers(s0, s1))
            if all(abs(p - x) >= 12 for x in cpos):
                r = ref[p]
                cpos.append(p)
                crl.append(1)
                calt.append(np.array([BASES[(int(np.flatnonzero(BASES == r)[0]) + rng.integers(1, 4)) % 4]],
                                     np.uint8))
                cph.append(-1)
        while len(cpos) > kmax:              # drop candidates first, then the last true site
            drop = max((i for i in range(len(cpos)) if cph[i] < 0), default=len(cpos) - 1)
            for lst in (cpos, crl, calt, cph):
                lst.pop(drop)
        o2 = np.argsort(cpos, kind="stable")
        cpos = [cpos[i] for i in o2]
        crl = [crl[i] for i in o2]
        calt = [calt[i] for i in o2]
        cph = [cph[i] for i in o2]
        k = len(cpos)
        for mask in range(1 << k):
            sel = [i for i in range(k) if mask >> i & 1]
            hseq = _apply_sites(ref, s0 - padding, s1 + padding, [cpos[i] for i in sel], [crl[i] for i in sel],
                                [calt[i] for i in sel])
            hb.append(hseq)
            hlen.append(len(hseq))
        nh.append(1 << k)
        m0 = sum(1 << i for i in range(k) if cph[i] == 0)
        m1 = sum(1 << i for i in range(k) if cph[i] == 1)
        a, b = min(m0, m1), max(m0, m1)
        true_pair.append(b * (b + 1) // 2 + a)
    hap_bytes = np.concatenate(hb) if hb else np.zeros(0, np.uint8)
    hap_off = np.concatenate([[0], np.cumsum(hlen)]).astype(np.int64)
    region_haps = np.concatenate([[0], np.cumsum(nh)]).astype(np.int64)
    batch = RegionBatch(read_bytes, qual_bytes, read_off, hap_bytes, hap_off, region_reads, region_haps,
                        gap=gap, ext=ext)
    info = {"seed": seed, "bases": L, "depth": depth, "read_len": read_len, "window": window,
            "padding": padding, "het_snvs": int(np.sum(kind == 0)), "het_indels": int(np.sum(kind == 1)),
            "sample_reads": int(nreads_total), "regions": int(nwin), "batch_reads": int(len(ridx)),
            "haplotypes": int(region_haps[-1]), "max_haps": int(max(nh) if nh else 0)}
    return ChromosomeSample(batch, mq, np.array(true_pair, np.int64), wstart, info)
