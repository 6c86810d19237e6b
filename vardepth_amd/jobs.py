"""Interned job lists: the host-side form of (read, quality, haplotype) requests.

The reference's likelihood functions receive Python lists of
``(sequence, quality, haplotype)`` triples (cpu_batch.py:20, trie_vector.py:16,
accelerator.py:166).  Every host path of this package turns such a list into
a :class:`JobTable` once:

* each distinct sequence, quality string and haplotype is stored once, as
  bytes, and jobs refer to them by small integers;
* deduplication, grouping by read and by haplotype set, and the scatter of
  device results back to job order are then numpy work on integer arrays;
* device batches (:class:`~vardepth_amd.workloads.RegionBatch`) and the
  ``char**`` arrays of the C-ABI drop-ins are assembled from the interned
  bytes (:class:`Packed`), so no string is encoded twice in one call.

Qualities that are not ``bytes`` (lists, integer arrays) are converted once per
object; the job list holds those objects for the whole call, so an object's
identity cannot be reused while it serves as a key.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .workloads import RegionBatch


def seq_bytes(s) -> bytes:
    """A read or haplotype as bytes (str must be ASCII, as the reference's encode('ascii'))."""
    return s.encode("ascii") if isinstance(s, str) else bytes(s)


def quality_bytes(q) -> bytes:
    return q if isinstance(q, bytes) else np.ascontiguousarray(q, dtype=np.uint8).tobytes()


class _Interner:
    """Value -> dense id, remembering the first object given for each value."""

    __slots__ = ("index", "values", "objects")

    def __init__(self):
        self.index = {}
        self.values = []
        self.objects = []

    def add(self, value, obj) -> int:
        i = self.index.get(value)
        if i is None:
            i = self.index[value] = len(self.values)
            self.values.append(value)
            self.objects.append(obj)
        return i


class JobTable:
    """Integer form of a job list.

    ``seq[j]``, ``qual[j]``, ``hap[j]`` are the ids of job j's read, quality
    and haplotype; ``seqs`` / ``quals`` / ``haps`` hold the interned bytes and
    ``seq_objects`` / ``hap_objects`` the objects the caller passed first for
    each value (used where a cache key must carry the caller's strings)."""

    def __init__(self, jobs, length_error: str = "Read and quality lengths differ"):
        count = len(jobs)
        self.count = count
        self.seq = np.empty(count, np.int64)
        self.qual = np.empty(count, np.int64)
        self.hap = np.empty(count, np.int64)
        seqs, quals, haps = _Interner(), _Interner(), _Interner()
        converted = {}   # id(quality object) -> bytes, for non-bytes qualities
        for j, (s, q, h) in enumerate(jobs):
            if len(s) != len(q):
                raise ValueError(length_error)
            if isinstance(q, bytes):
                qb = q
            else:
                qb = converted.get(id(q))
                if qb is None:
                    qb = converted[id(q)] = quality_bytes(q)
            self.seq[j] = seqs.add(s, s)
            self.qual[j] = quals.add(qb, qb)
            self.hap[j] = haps.add(h, h)
        self.seq_objects = seqs.objects
        self.hap_objects = haps.objects
        self.seqs = [seq_bytes(s) for s in seqs.values]
        self.quals = quals.values
        self.haps = [seq_bytes(h) for h in haps.values]

    # -- keys -----------------------------------------------------------
    def read_key(self) -> np.ndarray:
        """One int64 per job naming its (sequence, quality) read."""
        return self.seq * max(1, len(self.quals)) + self.qual

    def unique(self):
        """(first job of every distinct triple in first-occurrence order, job -> triple index)."""
        nq, nh = max(1, len(self.quals)), max(1, len(self.haps))
        if len(self.seqs) * nq * nh < 2 ** 62:
            key = (self.seq * nq + self.qual) * nh + self.hap
            _, first, inverse = np.unique(key, return_index=True, return_inverse=True)
        else:
            _, first, inverse = np.unique(np.stack([self.seq, self.qual, self.hap], 1), axis=0,
                                          return_index=True, return_inverse=True)
        inverse = inverse.reshape(-1)
        # renumber the triples by first occurrence, so that callers see input order
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(len(order))
        return first[order], rank[inverse]

    def job(self, j):
        """Job j as (caller's sequence, quality bytes, caller's haplotype)."""
        return self.seq_objects[self.seq[j]], self.quals[self.qual[j]], self.hap_objects[self.hap[j]]

    # -- device batches -------------------------------------------------
    def region_batch(self, read_jobs, hap_ids, gap=0.001, ext=0.1) -> RegionBatch:
        """A RegionBatch whose region k has the reads of the jobs listed in
        ``read_jobs[k]`` (one representative job per read) and the interned
        haplotypes ``hap_ids[k]``, in the given orders."""
        rjobs = np.asarray([j for js in read_jobs for j in js], np.int64)
        hids = np.asarray([h for hs in hap_ids for h in hs], np.int64)
        reads = Packed(self.seqs).take(self.seq[rjobs]) if len(rjobs) else (np.zeros(0, np.uint8), np.zeros(1, np.int64))
        quals = Packed(self.quals).take(self.qual[rjobs]) if len(rjobs) else (np.zeros(0, np.uint8), None)
        haps = Packed(self.haps).take(hids) if len(hids) else (np.zeros(0, np.uint8), np.zeros(1, np.int64))
        region_reads = np.zeros(len(read_jobs) + 1, np.int64)
        np.cumsum([len(js) for js in read_jobs], out=region_reads[1:])
        region_haps = np.zeros(len(hap_ids) + 1, np.int64)
        np.cumsum([len(hs) for hs in hap_ids], out=region_haps[1:])
        return RegionBatch(read_bytes=reads[0], qual_bytes=quals[0], read_off=reads[1], hap_bytes=haps[0],
                           hap_off=haps[1], region_reads=region_reads, region_haps=region_haps, gap=gap, ext=ext)


class Packed:
    """A list of byte strings as one contiguous uint8 buffer with offsets."""

    def __init__(self, items):
        self.lengths = np.fromiter((len(b) for b in items), np.int64, len(items))
        self.offsets = np.zeros(len(items) + 1, np.int64)
        np.cumsum(self.lengths, out=self.offsets[1:])
        self.buffer = np.frombuffer(b"".join(items), np.uint8) if len(items) else np.zeros(0, np.uint8)
        if not self.buffer.size:
            self.buffer = np.zeros(1, np.uint8)   # a valid address for empty strings

    def take(self, ids):
        """(bytes, offsets) of the strings ``ids`` concatenated in that order."""
        ids = np.asarray(ids, np.int64)
        lens = self.lengths[ids]
        off = np.zeros(len(ids) + 1, np.int64)
        np.cumsum(lens, out=off[1:])
        total = int(off[-1])
        if total == 0:
            return np.zeros(0, np.uint8), off
        # gather every byte: position p of output string k comes from offsets[ids[k]] + p
        starts = np.repeat(self.offsets[ids] - off[:-1], lens)
        return self.buffer[np.arange(total, dtype=np.int64) + starts], off

    def pointers(self, ids) -> ctypes.Array:
        """A ``char*`` array addressing strings ``ids`` inside the buffer
        (valid while this object lives)."""
        addr = (self.buffer.ctypes.data + self.offsets[np.asarray(ids, np.int64)]).astype(np.uint64)
        arr = (ctypes.c_char_p * len(addr))()
        if len(addr):
            ctypes.memmove(arr, addr.ctypes.data, addr.nbytes)
        return arr
