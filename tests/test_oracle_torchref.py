"""Pin the float64 autograd restatement (oracle/torch_ref.py) that every
gradient test uses as its reference.

* Values: the golden vectors written by the reference's own C++ compiled from
  source (tests/golden/make_golden.py): the shape grid in modes 0/1/2
  (pairhmm_pairs, batch_hmm.cpp:192) and the parametric g x x set through the
  scalar entry points (core.cpp:9 mode 0, anchored.cpp:5 mode 1, wide.cpp:9
  mode 2).
* Gradients, three independent ways:
  - exact rational derivatives of the brute-force path sum (oracle/brute.py)
    for n, m <= 4: the likelihood is a polynomial in each of g, x and every
    error bin, so sampling it exactly at more points than its degree and
    solving the Vandermonde system gives the exact derivative;
  - central differences (relative step 1e-3) of the C oracle (oracle.c,
    pinned bit for bit to the reference) on 150 bp reads of the golden region,
    in modes 0 and 1;
  - closed forms: mode 1 with n = m = 1 and equal bases gives
    L = (1-e)(1-2g); mode 0 with n = 1 gives
    L = (1-2g)/(m+1) sum_j eps_j + g/4.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest
import torch

import oracle
from oracle import brute, torch_ref
from tests.golden import fixtures as fx

G = fx.load()
ERR = oracle.error_table()[:64].copy()


def tvals(jobs, mode, gap=0.001, ext=0.1, err=ERR):
    e = torch.tensor(err, dtype=torch.float64)
    with torch.no_grad():
        return np.array([float(torch_ref.loglik(s, q, h, mode, torch.tensor(gap, dtype=torch.float64),
                                                torch.tensor(ext, dtype=torch.float64), e))
                         for s, q, h in jobs])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_values_match_golden_grid(mode):
    # shapes (1,1) .. (300,450); the (700,750) set is covered by the region test below
    jobs = fx.pairs("grid")[:30]
    got = tvals(jobs, mode)
    expect = G["grid_dense_m%d" % mode][:30]
    assert np.array_equal(np.isneginf(got), np.isneginf(expect))
    fin = np.isfinite(expect)
    np.testing.assert_allclose(got[fin], expect[fin], rtol=0, atol=1e-11)


@pytest.mark.parametrize("mode,key", [(0, "core"), (1, "anchored"), (2, "wide")])
def test_values_match_golden_parametric(mode, key):
    jobs = fx.pairs("param")
    for gi, g in enumerate(G["param_gaps"]):
        for xi, x in enumerate(G["param_exts"]):
            np.testing.assert_allclose(tvals(jobs, mode, float(g), float(x)),
                                       G["param_%s_g%d_x%d" % (key, gi, xi)], rtol=0, atol=1e-11)


@pytest.mark.parametrize("mode", [0, 1])
def test_values_match_golden_region(mode):
    jobs = fx.pairs("region")[::37]
    np.testing.assert_allclose(tvals(jobs, mode), G["region_dense_m%d" % mode][::37], rtol=0, atol=1e-11)


# --- exact derivatives of the brute-force likelihood --------------------------------------

def _solve(A, b):
    """Exact Gaussian elimination over Fractions."""
    n = len(b)
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [a - f * bb for a, bb in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def exact_derivative(fn, t0: Fraction, degree: int, step: Fraction) -> Fraction:
    """d/dt of the polynomial fn (degree <= degree) at t0, exactly."""
    pts = [t0 + (k - degree // 2) * step for k in range(degree + 1)]
    vals = [fn(t) for t in pts]
    coef = _solve([[t ** i for i in range(degree + 1)] for t in pts], vals)
    return sum(i * c * t0 ** (i - 1) for i, c in enumerate(coef) if i)


def _tiny_pairs():
    rng = np.random.default_rng(2024)
    out = []
    for n, m in [(1, 1), (2, 3), (3, 2), (3, 4), (4, 4), (4, 3)]:
        read = bytes(rng.choice(list(b"ACGT"), n).astype(np.uint8))
        hap = bytes(rng.choice(list(b"ACGT"), m).astype(np.uint8))
        if n <= m:
            hap = hap[:m - n] + read  # share bases so matches and mismatches both occur
        qual = bytes([int(v) for v in rng.choice([3, 10, 17, 30], n)])
        out.append((read, qual, hap))
    return out


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_grad_exact_against_brute_force(mode):
    gap, ext = Fraction(1, 64), Fraction(3, 16)
    # exact binary-fraction error table so float64 and Fraction see the same numbers
    err = [Fraction(1, 2 ** (1 + (q % 20))) + Fraction(1, 4096) for q in range(64)]
    errf = np.array([float(v) for v in err])
    for read, qual, hap in _tiny_pairs():
        n, m = len(read), len(hap)
        deg = n + m + 2

        def L(g=gap, x=ext, table=err):
            return brute.likelihood(read, [table[min(c, 63)] for c in qual], hap, g, x, mode)

        L0 = L()
        dg = exact_derivative(lambda t: L(g=t), gap, deg, Fraction(1, 256)) / L0
        dx = exact_derivative(lambda t: L(x=t), ext, deg, Fraction(1, 64)) / L0
        de = {}
        for b in sorted({min(c, 63) for c in qual}):
            def Lb(t, b=b):
                tab = list(err)
                tab[b] = t
                return L(table=tab)
            de[b] = exact_derivative(Lb, err[b], n + 1, Fraction(1, 8192)) / L0
        v, tg, tx, te = torch_ref.value_and_grad([(read, qual, hap)], mode, float(gap), float(ext), errf)
        assert math.isclose(v[0], math.log(L0), rel_tol=0, abs_tol=1e-13), (read, hap, v[0], math.log(L0))
        assert math.isclose(tg, float(dg), rel_tol=1e-11, abs_tol=1e-12), (read, hap, tg, float(dg))
        assert math.isclose(tx, float(dx), rel_tol=1e-11, abs_tol=1e-12), (read, hap, tx, float(dx))
        for b in range(64):
            expect = float(de.get(b, 0))
            assert math.isclose(te[b], expect, rel_tol=1e-11, abs_tol=1e-12), (read, hap, b, te[b], expect)


# --- central differences of the C oracle ----------------------------------------------------

@pytest.mark.parametrize("mode", [0, 1])
def test_grad_central_differences_of_c_oracle(mode):
    jobs = fx.pairs("region")[:40:5]
    gap, ext = 0.001, 0.1
    err = ERR.copy()
    _, tg, tx, te = torch_ref.value_and_grad(jobs, mode, gap, ext, err)

    def F(gp, xp, ep):
        e256 = np.array([ep[min(q, 63)] for q in range(256)])
        return sum(oracle.dense(s, q, h, mode, gp, xp, err=e256) for s, q, h in jobs)

    h = 1e-3
    fd_g = (F(gap * (1 + h), ext, err) - F(gap * (1 - h), ext, err)) / (2 * h * gap)
    fd_x = (F(gap, ext * (1 + h), err) - F(gap, ext * (1 - h), err)) / (2 * h * ext)
    assert abs(tg - fd_g) <= 1e-5 * abs(fd_g), (tg, fd_g)
    assert abs(tx - fd_x) <= 1e-5 * abs(fd_x), (tx, fd_x)
    used = sorted({min(c, 63) for _, q, _ in jobs for c in q})
    fd = {}
    for b in used:
        ep, em = err.copy(), err.copy()
        ep[b] *= 1 + h
        em[b] *= 1 - h
        fd[b] = (F(gap, ext, ep) - F(gap, ext, em)) / (2 * h * err[b])
    scale = max(abs(v) for v in fd.values())
    for b, v in fd.items():
        assert abs(te[b] - v) <= 1e-5 * scale, (b, te[b], v)
    assert np.all(te[[b for b in range(64) if b not in fd]] == 0)


# --- closed forms ---------------------------------------------------------------------------

def test_grad_closed_form_anchored_single_base():
    gap, ext = 0.004, 0.25
    err = ERR.copy()
    for q in (1, 7, 30, 63):
        v, tg, tx, te = torch_ref.value_and_grad([(b"C", bytes([q]), b"C")], 1, gap, ext, err)
        e = err[q]
        assert math.isclose(v[0], math.log((1 - e) * (1 - 2 * gap)), rel_tol=1e-14)
        assert math.isclose(tg, -2 / (1 - 2 * gap), rel_tol=1e-13)
        assert abs(tx) < 1e-15
        assert math.isclose(te[q], -1 / (1 - e), rel_tol=1e-13)


def test_grad_closed_form_semiglobal_one_row():
    gap, ext = 0.02, 0.3
    err = ERR.copy()
    q = 12
    e = err[q]
    hap = b"ACGTTGCA"
    for base in b"ACGT":
        m = len(hap)
        nmatch = sum(1 for c in hap if c == base)
        s_eps = nmatch * (1 - e) + (m - nmatch) * e / 3
        L = (1 - 2 * gap) / (m + 1) * s_eps + gap / 4
        dLdg = -2 / (m + 1) * s_eps + 0.25
        dLde = (1 - 2 * gap) / (m + 1) * (-nmatch + (m - nmatch) / 3)
        v, tg, tx, te = torch_ref.value_and_grad([(bytes([base]), bytes([q]), hap)], 0, gap, ext, err)
        assert math.isclose(v[0], math.log(L), rel_tol=1e-14)
        assert math.isclose(tg, dLdg / L, rel_tol=1e-12)
        assert abs(tx) < 1e-15
        assert math.isclose(te[q], dLde / L, rel_tol=1e-12, abs_tol=1e-14)
