"""Pin the genotype oracle (oracle/genotype.c): against the reference's own
mixture_general.cpp / mixture_two.cpp outputs (golden vectors), hand-computed
known answers, and invariants of the mixture that a wrong weight, index,
sign or dropped term breaks."""
import math

import numpy as np
import pytest

import oracle
from tests.golden import fixtures

# sum of ~100 log terms; the reference logs through libmvec, the oracle through libm
TOL = 1e-13


def _close(a, b, tol=TOL):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= tol * np.maximum(1.0, np.abs(b)))


def test_mixture_general_matches_reference_golden():
    cases = fixtures.mixture_general_cases()
    assert len(cases) > 80
    for m, prior, tri, expected in cases:
        got = oracle.mixture_general(m, prior, tri)
        assert got.shape == expected.shape
        assert _close(got, expected), (m.shape, prior, tri, np.max(np.abs(got - expected)))


def test_mixture_two_matches_reference_golden():
    for m, prior, expected in fixtures.mixture_two_cases():
        got = oracle.mixture_two(m, prior)
        assert _close(got, expected), (m.shape, prior)


def test_mixture_two_equals_general_square():
    """mixture_two.cpp and mixture_general.cpp state the same model for n = 2."""
    rng = np.random.default_rng(3)
    m = rng.random((30, 2)) ** 2
    for prior in (0.0, 1e-3, 0.4):
        assert _close(oracle.mixture_two(m, prior).ravel(), oracle.mixture_general(m, prior, 0))


def test_hand_computed_single_row():
    """lh = [[1, 0]]: pair (0,0) -> 0; (0,1) -> log(sum bw_j w_j) = log .5;
    (1,1) -> log(1e-300).  With the artifact prior p: (0,0) -> log((1-p) + p
    (1 - sum tw theta / 2)) = log(1 - .049 p); (0,1) -> log .5."""
    m = np.array([[1.0, 0.0]])
    got = oracle.mixture_general(m, 0.0, 1)
    assert got[0] == pytest.approx(0.0, abs=1e-15)
    assert got[1] == pytest.approx(math.log(0.5), abs=1e-15)
    assert got[2] == pytest.approx(math.log(1e-300), abs=1e-12)
    for p in (1e-3, 0.25):
        got = oracle.mixture_general(m, p, 1)
        assert got[0] == pytest.approx(math.log(1 - 0.049 * p), abs=1e-15)
        assert got[1] == pytest.approx(math.log(0.5), abs=1e-15)
    # no mixture: mean of the pair
    got = oracle.mixture_general(m, None, 1)
    assert got[1] == pytest.approx(math.log(0.5), abs=1e-15)


def test_one_haplotype_and_equal_columns_reduce_to_sum_of_logs():
    """Every mixture weight set sums to one, so when all columns are equal every
    pair's likelihood is sum_r log l_r, whatever the prior."""
    rng = np.random.default_rng(4)
    l = rng.random(25) * 0.9 + 0.05
    s = float(np.sum(np.log(l)))
    for n in (1, 3, 6):
        m = np.repeat(l[:, None], n, axis=1)
        for prior in (None, 0.0, 1e-3, 0.5):
            got = oracle.mixture_general(m, prior, 1)
            assert np.allclose(got, s, rtol=0, atol=1e-11), (n, prior)


def test_square_output_is_symmetric_and_triangular_is_its_lower_order():
    rng = np.random.default_rng(5)
    m = rng.random((40, 5)) ** 3
    for prior in (None, 0.0, 2e-3):
        sq = oracle.mixture_general(m, prior, 0).reshape(5, 5)
        assert np.allclose(sq, sq.T, rtol=0, atol=1e-12)
        tri = oracle.mixture_general(m, prior, 1)
        order = [(a, b) for b in range(5) for a in range(b + 1)]
        assert np.allclose(tri, [sq[a, b] for a, b in order], rtol=0, atol=1e-12)


def test_column_permutation_permutes_pairs():
    rng = np.random.default_rng(6)
    n = 6
    m = rng.random((35, n)) ** 2
    perm = rng.permutation(n)
    for prior in (0.0, 1e-3):
        sq = oracle.mixture_general(m, prior, 0).reshape(n, n)
        sqp = oracle.mixture_general(m[:, perm], prior, 0).reshape(n, n)
        assert np.allclose(sqp, sq[np.ix_(perm, perm)], rtol=0, atol=1e-11)


def test_row_scaling_adds_log_scale():
    """Every mixture value is linear in the row, so scaling a row by s adds log s
    to every pair (rows kept far above the 1e-300 floor)."""
    rng = np.random.default_rng(7)
    m = rng.random((20, 4)) * 0.9 + 0.1
    s = rng.random(20) * 4 + 0.5
    for prior in (None, 0.0, 0.01):
        a = oracle.mixture_general(m, prior, 1)
        b = oracle.mixture_general(m * s[:, None], prior, 1)
        assert np.allclose(b - a, np.sum(np.log(s)), rtol=0, atol=1e-11)


def test_rows_zero_and_row_order():
    m = np.random.default_rng(8).random((12, 3))
    z = oracle.mixture_general(m[:0], 1e-3, 1)
    assert np.allclose(z, 0.0, atol=1e-15)             # log of weights that sum to one
    a = oracle.mixture_general(m, 1e-3, 1)
    b = oracle.mixture_general(m[::-1].copy(), 1e-3, 1)
    assert np.allclose(a, b, rtol=0, atol=1e-12)


def test_normalized_rows_hand_case():
    """ll = [0, ln .5], mq 10: e = .1, exp -> [1, .5], mean .75,
    values = [.9 + .075, .45 + .075], certainty = ln(.975 / .525)."""
    v, c = oracle.normalized_rows(np.array([[0.0, math.log(0.5)]]), [10])
    assert v[0, 0] == pytest.approx(0.975, abs=1e-15)
    assert v[0, 1] == pytest.approx(0.525, abs=1e-15)
    assert c[0] == pytest.approx(math.log(0.975 / 0.525), abs=1e-14)


def test_normalized_rows_properties():
    rng = np.random.default_rng(9)
    ll = rng.normal(-50, 20, (30, 5))
    mq = rng.integers(0, 256, 30).astype(np.uint8)
    v, c = oracle.normalized_rows(ll, mq)
    v2, c2 = oracle.normalized_rows(ll + rng.normal(0, 100, (30, 1)), mq)   # shift per row
    assert np.allclose(v, v2, rtol=1e-12, atol=0)
    e = 10.0 ** (-np.minimum(mq.astype(np.int64), 60) / 10)
    mean_exp = np.exp(ll - ll.max(1, keepdims=True)).mean(1)
    assert np.allclose(v.max(axis=1), (1 - e) + e * mean_exp, rtol=1e-14, atol=0)   # the row max maps to 1
    hi, _ = oracle.normalized_rows(ll, np.full(30, 60, np.uint8))
    hi2, _ = oracle.normalized_rows(ll, np.full(30, 255, np.uint8))
    assert np.array_equal(hi, hi2)                      # mq clamps at 60
    z, cz = oracle.normalized_rows(ll, np.zeros(30, np.uint8))   # e = 1: every value is the mean
    assert np.allclose(z, z[:, :1]) and np.allclose(cz, 0.0)
    assert np.allclose(c, np.log(v.max(1)) - np.log(v.min(1)), rtol=0, atol=1e-12)
    nanrow = ll.copy()
    nanrow[3, 2] = np.nan
    vn, cn = oracle.normalized_rows(nanrow, mq)
    assert np.isnan(vn[3]).all() and np.isnan(cn[3]) and np.isfinite(cn[4])


def test_genotype_regions_composes_the_chain():
    """The batched chain equals normalized rows -> certainty filter -> molecule
    rows -> diploid likelihood applied region by region."""
    rng = np.random.default_rng(10)
    ncol = np.array([2, 1, 4, 3, 2, 6], np.int32)
    nrow = np.array([7, 3, 0, 12, 5, 20])
    row_off = np.concatenate([[0], np.cumsum(nrow)])
    mat_off = np.concatenate([[0], np.cumsum(nrow * ncol)])[:-1]
    ll = rng.normal(-30, 6, int(np.sum(nrow * ncol)))
    mq = rng.integers(0, 70, int(row_off[-1])).astype(np.uint8)
    gl_off = np.concatenate([[0], np.cumsum(ncol * (ncol + 1) // 2)])
    gl, bad = oracle.genotype_regions(row_off, ncol, mat_off, ll, mq, gl_off, threads=3)
    assert bad == 0
    for k in range(len(ncol)):
        n = int(ncol[k])
        block = ll[mat_off[k]:mat_off[k] + nrow[k] * n].reshape(nrow[k], n)
        v, c = oracle.normalized_rows(block, mq[row_off[k]:row_off[k + 1]])
        keep = ~(c < math.log(3))
        lv = np.log(np.maximum(1e-300, v[keep]))
        mol = np.exp(lv - lv.max(axis=1, keepdims=True)) if keep.any() else np.empty((0, n))
        exp = oracle.diploid_likelihood(mol, 1e-3)
        assert np.allclose(gl[gl_off[k]:gl_off[k + 1]], exp, rtol=0, atol=1e-12), k


def test_planted_mistake_is_caught():
    """Swapping two balance weights (a plausible transcription slip) moves a
    heterozygous pair's likelihood by far more than the test tolerance."""
    m = np.array([[0.9, 0.1], [0.2, 0.7], [0.6, 0.3]])
    good = oracle.mixture_general(m, 0.0, 1)
    bal = np.array([.2, .3, .4, .5, .6, .7, .8])
    bw = np.array([.025, .1, .225, .3, .225, .1, .025])
    bw_bad = bw.copy()
    bw_bad[[0, 1]] = bw_bad[[1, 0]]
    a, b = m[:, 0], m[:, 1]
    bad = math.log(sum(w * math.exp(np.sum(np.log(x * a + (1 - x) * b))) for x, w in zip(bal, bw_bad)))
    assert abs(bad - good[1]) > 1e-3


# ---- near-tie guard (partial_anchor.py:56-57,119-127; whole_read.py:51) -------------------

def _one_region(gl, rows_cert, n):
    return dict(row_off=[0, len(rows_cert)], ncol=[n], gl_off=[0], gl=np.asarray(gl, float),
                cert=np.asarray(rows_cert, float))


def test_guard_hand_cases_uniform_prior():
    # n = 2: pairs (0,0), (0,1), (1,1); uniform prior 1/4, 1/2, 1/4
    gl = [-10.0, -10.00005, -30.0]
    f, g = oracle.guard(**_one_region(gl, [5.0], 2))
    assert g[0, 0] == pytest.approx(5e-5, rel=1e-9)
    assert g[0, 1] == pytest.approx(math.log(2) - 5e-5, rel=1e-12)   # the het pair gains log 2
    assert f[0] == 1                                                  # 5e-5 <= 1e-4 * max(1, 1)
    f, _ = oracle.guard(**_one_region([-10.0, -10.0002, -30.0], [5.0], 2))
    assert f[0] == 0                                                  # 2e-4 > 1e-4
    # molecules = rows whose certainty is not below ln 3: the margin scales with them
    f, _ = oracle.guard(**_one_region([-10.0, -10.00025, -30.0], [5.0, 5.0, 0.5], 2))
    assert f[0] == 0                                                  # two molecules: 2.5e-4 > 2e-4
    f, _ = oracle.guard(**_one_region([-10.0, -10.00025, -30.0], [5.0, 5.0, 5.0], 2))
    assert f[0] == 1                                                  # three molecules: 2.5e-4 <= 3e-4
    f, _ = oracle.guard(**_one_region([-10.0, -10.00025, -30.0], [5.0, 5.0, float("nan")], 2))
    assert f[0] == 1                                                  # NaN certainty counts (not below ln 3)
    # the posterior gap alone can flag: gl clearly separated, the prior closes the gap
    gl = [-10.0, -10.0 - math.log(2), -40.0]
    f, g = oracle.guard(**_one_region(gl, [5.0], 2))
    assert g[0, 0] == pytest.approx(math.log(2)) and g[0, 1] == pytest.approx(0.0, abs=1e-14) and f[0] == 1
    # a single pair is never flagged
    f, g = oracle.guard(**_one_region([-3.0], [5.0], 1))
    assert f[0] == 0 and np.isinf(g).all()


def test_guard_allele_frequency_prior():
    n = 3
    pairs = [(a, b) for b in range(n) for a in range(b + 1)]
    # equal frequencies (including floored zeros) reproduce the uniform prior exactly
    rng = np.random.default_rng(3)
    gl = rng.normal(-20, 3, len(pairs))
    base = oracle.guard(**_one_region(gl, [5.0], n))
    for af in ([0.2, 0.2, 0.2], [0.0, 0.0, 0.0], [0.0005, 0.001, 0.0]):
        f, g = oracle.guard(**_one_region(gl, [5.0], n), af=af)
        np.testing.assert_allclose(g, base[1], rtol=0, atol=1e-12)
    # a constant gl: the posterior gap is the log ratio of the two largest prior weights
    af = np.array([0.7, 0.2, 0.1])
    p = np.array([af[a] * af[b] * (1 if a == b else 2) for a, b in pairs])
    top = np.sort(p)[::-1]
    f, g = oracle.guard(**_one_region(np.full(len(pairs), -5.0), [5.0], n), af=af)
    assert g[0, 0] == 0.0 and f[0] == 1
    assert g[0, 1] == pytest.approx(math.log(top[0] / top[1]), rel=1e-12)   # 0.49 vs 0.28
    # frequencies are floored at .001 before normalising (partial_anchor.py:56)
    f1, g1 = oracle.guard(**_one_region(gl, [5.0], n), af=[0.5, 0.5, 0.0])
    f2, g2 = oracle.guard(**_one_region(gl, [5.0], n), af=[0.5, 0.5, 0.001])
    np.testing.assert_array_equal(g1, g2)
    # scaling all frequencies leaves the prior unchanged
    f3, g3 = oracle.guard(**_one_region(gl, [5.0], n), af=[5.0, 5.0, 0.01])
    np.testing.assert_allclose(g3, g2, rtol=0, atol=1e-12)


def _numpy_form(lh, artifact_prior):
    """triangular.py:13-24 evaluated with numpy (the reference's canonical form)."""
    balance = (.2, .3, .4, .5, .6, .7, .8)
    bw = (.025, .1, .225, .3, .225, .1, .025)
    theta = (.02, .1, .2, .3)
    tw = (.4, .35, .2, .05)
    n = lh.shape[1]
    a, b = np.array([(a, b) for b in range(n) for a in range(b + 1)], dtype=np.int64).T
    la, lb = lh[:, a], lh[:, b]
    if artifact_prior is None:
        return np.log(np.maximum(1e-300, (la + lb) / 2)).sum(axis=0)
    clean = np.array([np.log(np.maximum(1e-300, w * la + (1 - w) * lb)).sum(axis=0) + math.log(p)
                      for w, p in zip(balance, bw)])
    clean = np.logaddexp.reduce(clean, axis=0)
    if artifact_prior:
        art = []
        for t, w in zip(theta, tw):
            for c in range(n):
                art.append(np.log(np.maximum(1e-300, (1 - t) * (la + lb) / 2 + t * lh[:, c, None])).sum(axis=0)
                           + math.log(w / n))
        art = np.logaddexp.reduce(np.array(art), axis=0)
        clean = np.logaddexp(math.log1p(-artifact_prior) + clean, math.log(artifact_prior) + art)
    return clean


def test_numpy_canonical_form_equals_mixture_general():
    """The reference's two forms of the diploid likelihood -- the numpy one of
    triangular.py:13-24 and mixture_general.cpp -- agree to the last bits of
    their logs, so the device mixture (pinned to mixture_general) serves both."""
    rng = np.random.default_rng(61)
    for n in (1, 2, 3, 7, 16):
        for rows in (0, 1, 9, 40):
            lh = rng.choice([1.0, 0.3, 1e-3, 1e-200, 0.0], (rows, n)) * rng.uniform(0.5, 1, (rows, n))
            for prior in (None, 0.0, 1e-3, 0.3):
                want = _numpy_form(lh, prior)
                got = oracle.mixture_general(lh, prior, triangular=True)
                np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-12)
