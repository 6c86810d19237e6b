"""Exact likelihood by explicit enumeration of alignments (TEST INFRASTRUCTURE).

This is the textbook definition of the pair-HMM likelihood -- a sum over
every alignment path of the product of its transition and emission
probabilities -- written without any dynamic-programming table, so that it
pins the recurrence of the oracle (and of the GPU engine) independently:
an off-by-one column, a swapped transition or a missing term changes the
set of paths or their weights.

Model (batch_hmm.cpp:14,30,45-51; anchored.cpp:8-23):

* states M (consume read + haplotype base), I (consume read base only),
  D (consume haplotype base only);
* M->M = 1-2g, M->I = M->D = g, I->I = D->D = x, I->M = D->M = 1-x,
  no I<->D transitions;
* M emits 1-e on equal bytes else e/3; I emits 1/4; D emits nothing;
* semiglobal (modes 0, 2): start in M "before" read base 1 at any haplotype
  column s in [0, m] with probability 1/(m+1); no deletion before the first
  read base; the path ends when the last read base is emitted, in M or I;
* anchored (mode 1): start in M at column 0 with probability 1; deletions
  before the first read base allowed; the path must end at column m after the
  last read base, in any state.

All arithmetic uses ``fractions.Fraction`` so the result is exact for exact
(binary-fraction) parameters.
"""
from __future__ import annotations

from fractions import Fraction


def likelihood(read: bytes, qual_err, hap: bytes, gap, ext, mode: int) -> Fraction:
    """Exact likelihood. ``qual_err[i]`` is the error probability of read base i."""
    n, m = len(read), len(hap)
    if n == 0 or m == 0:
        return Fraction(0)
    g, x = Fraction(gap), Fraction(ext)
    trans = {("M", "M"): 1 - 2 * g, ("M", "I"): g, ("M", "D"): g,
             ("I", "I"): x, ("I", "M"): 1 - x, ("D", "D"): x, ("D", "M"): 1 - x}
    quarter = Fraction(1, 4)
    total = Fraction(0)

    def walk(i, j, state, weight):
        # i read bases and j haplotype bases consumed; `state` emitted last
        nonlocal total
        if i == n:
            if mode == 1:
                if j == m:
                    total += weight
            elif state in ("M", "I"):
                total += weight
                return  # semiglobal paths end at the last read base
        for nxt in ("M", "I", "D"):
            t = trans.get((state, nxt))
            if t is None:
                continue
            if nxt == "M":
                if i == n or j == m:
                    continue
                e = Fraction(qual_err[i])
                emit = (1 - e) if read[i] == hap[j] else e / 3
                walk(i + 1, j + 1, "M", weight * t * emit)
            elif nxt == "I":
                if i == n:
                    continue
                walk(i + 1, j, "I", weight * t * quarter)
            else:
                if j == m:
                    continue
                if i == 0 and mode != 1:
                    continue  # row 0 carries no deletion state when semiglobal
                walk(i, j + 1, "D", weight * t)

    if mode == 1:
        walk(0, 0, "M", Fraction(1))
    else:
        for s in range(m + 1):
            walk(0, s, "M", Fraction(1, m + 1))
    return total
