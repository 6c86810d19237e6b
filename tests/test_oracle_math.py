"""Pin the oracle where the golden vectors leave it open: the exact sum over
alignments on tiny inputs, closed forms, and invariants of the model."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from oracle import brute

# binary-fraction parameters: exact in FP64, so the only FP64 error is rounding
GAP, EXT = 2.0 ** -7, 2.0 ** -3
ERR = np.array([2.0 ** -(1 + (q % 6)) for q in range(64)])   # 1/2 .. 1/64


def _exact(read, qual, hap, mode):
    e = [Fraction(ERR[min(q, 63)]) for q in qual]
    return brute.likelihood(read, e, hap, Fraction(GAP), Fraction(EXT), mode)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_dense_equals_sum_over_alignments(mode):
    rng = np.random.default_rng(7)
    for _ in range(60):
        n, m = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        read = rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes()
        hap = rng.choice(list(b"ACGT"), m).astype(np.uint8).tobytes()
        qual = bytes(rng.integers(0, 64, n).astype(np.uint8))
        exact = _exact(read, qual, hap, mode)
        got = oracle.dense(read, qual, hap, mode, GAP, EXT, ERR)
        assert got == pytest.approx(math.log(exact), abs=1e-12), (read, hap, qual)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_trie_equals_sum_over_alignments(mode):
    rng = np.random.default_rng(8)
    for _ in range(20):
        n = int(rng.integers(1, 5))
        read = rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes()
        qual = bytes(rng.integers(0, 64, n).astype(np.uint8))
        haps = [rng.choice(list(b"ACGT"), int(rng.integers(1, 5))).astype(np.uint8).tobytes() for _ in range(3)]
        out, _ = oracle.trie_many([read], [qual], haps, mode, 0, GAP, EXT, ERR)
        for j, h in enumerate(haps):
            assert out[0, j] == pytest.approx(math.log(_exact(read, qual, h, mode)), abs=1e-12)


def test_brute_force_detects_planted_mistakes():
    """A dropped gap-extension term or a wrong column index must change the value."""
    read, hap, qual = b"ACG", b"AGCG", bytes([3, 7, 11])
    good = _exact(read, qual, hap, 0)
    # planted mistake 1: no gap extension (drops every x term)
    e = [Fraction(ERR[q]) for q in qual]
    no_del = brute.likelihood(read, e, hap, Fraction(GAP), Fraction(0), 0)
    assert no_del != good
    # planted mistake 2: shifted haplotype (index error) changes the likelihood
    assert _exact(read, qual, hap[1:] + hap[:1], 0) != good
    # and the oracle tracks the true value, not the planted ones
    got = oracle.dense(read, qual, hap, 0, GAP, EXT, ERR)
    assert got == pytest.approx(math.log(good), abs=1e-13)
    assert abs(got - math.log(no_del)) > 1e-6


def test_closed_forms():
    # anchored n=m=1, equal bases: L = (1-e)(1-2g)
    e = oracle.error_table()[30]
    v = oracle.dense(b"A", bytes([30]), b"A", 1)
    assert v == pytest.approx(math.log((1 - e) * (1 - 2 * 0.001)), abs=1e-15)
    # semiglobal n=1: L = (1-2g)/(m+1) * sum_j emit_j + g/4
    hap = b"ACGTTA"
    m = len(hap)
    emits = [(1 - e) if c == ord("A") else e / 3 for c in hap]
    expect = (1 - 2 * 0.001) / (m + 1) * sum(emits) + 0.001 / 4
    for mode in (0, 2):
        assert oracle.dense(b"A", bytes([30]), hap, mode) == pytest.approx(math.log(expect), abs=1e-14)


def test_invariants():
    rng = np.random.default_rng(99)
    for _ in range(30):
        n, m = int(rng.integers(1, 120)), int(rng.integers(1, 160))
        read = rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes()
        hap = rng.choice(list(b"ACGT"), m).astype(np.uint8).tobytes()
        qual = bytes(rng.integers(0, 61, n).astype(np.uint8))
        v0 = oracle.dense(read, qual, hap, 0)
        v2 = oracle.dense(read, qual, hap, 2)
        v1 = oracle.dense(read, qual, hap, 1)
        assert v0 == pytest.approx(v2, abs=1e-9)          # modes 0 and 2 are the same model
        assert v0 <= 1e-12 and v1 <= 1e-12               # likelihoods of a read are <= 1
    # anchored(read, read) beats a 1-bp indel variant by > 3 nats (test_v03.py:76-81)
    for read, wrong in [(b"ACGTTACG", b"ACGTACG"), (b"ACGTACG", b"ACGTTACG")]:
        q = bytes([40] * len(read))
        assert oracle.dense(read, q, read, 1, strict=True) > oracle.dense(read, q, wrong, 1, strict=True) + 3


def test_error_table():
    e = oracle.error_table()
    assert e[0] == 1.0 and e[10] == pytest.approx(0.1) and e[60] == e[61] == e[255]
    np.testing.assert_array_equal(e, [math.pow(10.0, -min(60, q) / 10.0) for q in range(256)])
