"""Device diploid genotype likelihood vs the oracle (needs an MI355X).

The device forms every row's mixture value with the reference's operation
order and no contraction, and carries each total as an exact-exponent product
logged once; the oracle sums logs row by row as the reference does.  They
agree to the rounding of that sum: TOL relative per output (1e-12 of a total
that is a sum of ~100-1000 log terms; the observed gap is ~1e-15).  Pair
indices (triangular order b = 0..n-1, a = 0..b) are compared exactly: every
output is checked at its own index."""
import math

import numpy as np
import pytest

import oracle
from tests.golden import fixtures
from vardepth_amd import engine, workloads, _lib
from vardepth_amd import mixture_cpu, molecule_rows, triangular

pytestmark = pytest.mark.gpu
TOL = 1e-12


def close(a, b, tol=TOL):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    err = np.abs(a[ok] - b[ok]) / np.maximum(1.0, np.abs(b[ok]))
    assert err.size == 0 or err.max() <= tol, (err.max(), np.argmax(err))
    return float(err.max()) if err.size else 0.0


def test_mixture_general_matches_reference_golden():
    for m, prior, tri, expected in fixtures.mixture_general_cases():
        got = mixture_cpu.general(m, prior, triangular=bool(tri))
        close(np.ravel(got), expected)


def test_mixture_two_matches_reference_golden():
    for m, prior, expected in fixtures.mixture_two_cases():
        close(mixture_cpu.two(m, prior), expected)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 9, 16, 33, 64])
def test_mixture_general_random_vs_oracle(n):
    rng = np.random.default_rng(100 + n)
    for rows in ((0, 1, 5, 63, 64, 65, 300) if n <= 16 else (0, 3, 65)):
        m = rng.random((rows, n)) ** 3
        if rows > 2:
            m[1, 0] = 0.0
        for prior in (None, 0.0, 1e-3, 0.3):
            for tri in ((True, False) if n <= 16 else (True,)):
                got = mixture_cpu.general(m, prior, triangular=tri)
                close(np.ravel(got), oracle.mixture_general(m, prior, tri))


def test_diploid_likelihood_dispatch_and_order():
    rng = np.random.default_rng(5)
    m2 = rng.random((40, 2))
    close(triangular.diploid_likelihood(m2, 1e-3), oracle.diploid_likelihood(m2, 1e-3))
    close(triangular.diploid_likelihood(m2), oracle.diploid_likelihood(m2))
    # a matrix whose only well-supported genotype is (a, b) = (1, 3): its index is 3*4/2 + 1 = 7
    n = 5
    m = np.full((60, n), 1e-6)
    m[:30, 1] = 1.0
    m[30:, 3] = 1.0
    gl = triangular.diploid_likelihood(m, 1e-3)
    assert int(np.argmax(gl)) == 7
    assert triangular.pair_order(n)[7] == (1, 3)


def test_validation_matches_reference():
    with pytest.raises(ValueError):
        mixture_cpu.general(np.array([[0.5, np.nan]]), None)
    with pytest.raises(ValueError):
        mixture_cpu.general(np.array([[0.5, -1.0]]), 0.0)
    with pytest.raises(ValueError):
        mixture_cpu.general(np.array([[0.5, 0.2]]), 1.0)
    with pytest.raises(ValueError):
        mixture_cpu.two(np.array([[0.5, 0.2]]), -0.1)
    with pytest.raises(ValueError):
        molecule_rows.normalized_rows(np.zeros((3, 2)), [60, 60])
    # the C ABI rejects the same inputs on its own
    L = _lib.lib()
    bad = np.array([0.5, np.inf])
    out = np.empty(3)
    dp = lambda a: a.ctypes.data_as(_lib._dp)
    assert L.vd_mixture_general(1, 2, dp(bad), -1.0, 1, dp(out)) != 0
    ok = np.array([0.5, 0.25])
    assert L.vd_mixture_general(1, 2, dp(ok), 1.5, 1, dp(out)) != 0


def test_normalized_rows_vs_oracle():
    rng = np.random.default_rng(11)
    for n in (1, 2, 5, 64):
        ll = rng.normal(-40, 25, (150, n))
        ll[3, 0] = -np.inf
        mq = rng.integers(0, 256, 150).astype(np.uint8)
        v, c = molecule_rows.normalized_rows(ll, mq)
        ve, ce = oracle.normalized_rows(ll, mq)
        close(v, ve, 1e-14)
        close(c, ce, 1e-12)
    ll = rng.normal(-10, 3, (4, 3))
    ll[1, 2] = np.nan
    v, c = molecule_rows.normalized_rows(ll, [60, 60, 60, 60])
    assert np.isnan(v[1]).all() and np.isnan(c[1]) and np.isfinite(c[[0, 2, 3]]).all()
    e, ce = molecule_rows.normalized_rows(np.empty((0, 3)), [])
    assert e.shape == (0, 3) and ce.shape == (0,)


def _chain_oracle(b, ll, mq, min_cert=math.log(3), prior=1e-3):
    nh = np.diff(b.region_haps).astype(np.int32)
    gl_off = engine.gl_offsets(nh)
    gl, bad = oracle.genotype_regions(b.region_reads, nh, b.out_off()[:-1], ll, mq, gl_off, min_cert, prior)
    return gl, bad, gl_off


@pytest.fixture(scope="module")
def sample():
    return workloads.chromosome_sample(mb=0.3, depth=30.0)


def test_genotype_chain_on_chromosome_sample(sample):
    import torch
    b = sample.batch
    dev = engine.DeviceBatch(b)
    dev.forward(mode=0)
    mq = torch.from_numpy(sample.mq).cuda()
    gl, vals, cert, st = dev.genotype(mq, values=True, stats=True)
    torch.cuda.synchronize()
    ll = dev.result()
    exp, bad, gl_off = _chain_oracle(b, ll, sample.mq)
    assert bad == 0 and st["invalid"] == 0 and st["skipped"] == 0
    close(gl.cpu().numpy(), exp)
    # values / certainty of the first stage, and the kept-row count
    v = vals.cpu().numpy()
    c = cert.cpu().numpy()
    oo = b.out_off()
    kept = 0
    for k in range(0, b.nregions, 7):
        r0, r1 = b.region_reads[k], b.region_reads[k + 1]
        n = b.region_haps[k + 1] - b.region_haps[k]
        block = ll[oo[k]:oo[k + 1]].reshape(r1 - r0, n)
        ve, ce = oracle.normalized_rows(block, sample.mq[r0:r1])
        close(v[oo[k]:oo[k + 1]].reshape(r1 - r0, n), ve, 1e-14)
        close(c[r0:r1], ce, 1e-12)
    kept = int(np.sum(~(c < math.log(3))))
    assert st["kept_rows"] == kept
    # the genotype calls recover the simulated truth in nearly every variant window
    nh = np.diff(b.region_haps)
    g = gl.cpu().numpy()
    calls = [int(np.argmax(g[gl_off[k]:gl_off[k + 1]])) for k in range(b.nregions)]
    var = nh >= 2
    assert np.mean(np.array(calls)[var] == sample.true_pair[var]) > 0.97


def test_genotype_edge_regions():
    """Empty regions, a region with no rows, one haplotype, every row filtered,
    a NaN likelihood row (kept by the filter, so the region is invalid), many
    rows per lane, 64 haplotypes."""
    import torch
    rng = np.random.default_rng(21)
    shapes = [(0, 3), (5, 1), (12, 2), (0, 0), (70, 5), (130, 64), (9, 4), (3, 3)]
    blocks, ncols, nrows = [], [], []
    for rows, n in shapes:
        blk = rng.normal(-60, 15, (rows, n))
        blocks.append(blk)
        ncols.append(n)
        nrows.append(rows)
    blocks[6][:] = -5.0            # equal columns: certainty 0, every row dropped
    blocks[7][1, 1] = np.nan       # NaN row: kept, region invalid
    ncols, nrows = np.array(ncols), np.array(nrows)
    row_off = np.concatenate([[0], np.cumsum(nrows)]).astype(np.int64)
    col_off = np.concatenate([[0], np.cumsum(ncols)]).astype(np.int64)
    mat_off = np.concatenate([[0], np.cumsum(nrows * ncols)])[:-1].astype(np.int64)
    ll = np.concatenate([b.ravel() for b in blocks])
    mq = rng.integers(10, 70, int(row_off[-1])).astype(np.uint8)
    gl_off = engine.gl_offsets(ncols)
    exp, bad = oracle.genotype_regions(row_off, ncols.astype(np.int32), mat_off, ll, mq, gl_off)
    assert bad == 1
    # NULL stream: the default stream, ordered after torch's copies above; stats synchronises
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    d_row, d_col, d_mat, d_gl = t(row_off), t(col_off), t(mat_off), t(gl_off)
    d_ll, d_mq = t(ll), t(mq)
    out = torch.full((int(gl_off[-1]),), -7.0, dtype=torch.float64, device="cuda")
    gb = _lib.VdGenoBatch(d_row.data_ptr(), d_col.data_ptr(), d_mat.data_ptr(), d_gl.data_ptr(), len(ncols),
                          int(row_off[-1]), len(ll), int(gl_off[-1]))
    st = _lib.VdGenoStats()
    import ctypes
    code = _lib.lib().vd_genotype_device(ctypes.byref(gb), ctypes.c_void_p(d_ll.data_ptr()),
                                         ctypes.c_void_p(d_mq.data_ptr()), math.log(3), 1e-3,
                                         ctypes.c_void_p(out.data_ptr()), None, None, None, ctypes.byref(st))
    _lib.check(code, "vd_genotype_device")
    assert st.invalid == 1 and st.skipped == 0
    close(out.cpu().numpy(), exp)


def test_mixture_regions_device_square_with_keep():
    import ctypes
    import torch
    rng = np.random.default_rng(31)
    shapes = [(20, 3), (1, 1), (90, 9), (0, 4), (40, 2)]
    mats = [rng.random(s) ** 2 for s in shapes]
    keeps = [rng.random(s[0]) < 0.7 for s in shapes]
    ncols = np.array([s[1] for s in shapes])
    nrows = np.array([s[0] for s in shapes])
    row_off = np.concatenate([[0], np.cumsum(nrows)]).astype(np.int64)
    col_off = np.concatenate([[0], np.cumsum(ncols)]).astype(np.int64)
    mat_off = np.concatenate([[0], np.cumsum(nrows * ncols)])[:-1].astype(np.int64)
    for tri in (0, 1):
        gl_off = engine.gl_offsets(ncols, triangular=bool(tri))
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        d = [t(row_off), t(col_off), t(mat_off), t(gl_off)]
        lh = t(np.concatenate([m.ravel() for m in mats]))
        keep = t(np.concatenate(keeps).astype(np.int8))
        out = torch.empty(int(gl_off[-1]), dtype=torch.float64, device="cuda")
        gb = _lib.VdGenoBatch(*[x.data_ptr() for x in d], len(shapes), int(row_off[-1]), lh.numel(),
                              int(gl_off[-1]))
        for prior in (-1.0, 0.0, 0.01):
            code = _lib.lib().vd_mixture_regions_device(ctypes.byref(gb), ctypes.c_void_p(lh.data_ptr()),
                                                        ctypes.c_void_p(keep.data_ptr()), prior, tri,
                                                        ctypes.c_void_p(out.data_ptr()),
                                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
                                                        None)
            _lib.check(code, "vd_mixture_regions_device")
            got = out.cpu().numpy()
            for k, (m, kp) in enumerate(zip(mats, keeps)):
                exp = oracle.mixture_general(m[kp], None if prior < 0 else prior, tri)
                close(got[gl_off[k]:gl_off[k + 1]], exp)


def test_genotype_graph_replay(sample):
    """Forward + genotype captured in one HIP graph: every replay equals the
    eager result."""
    import torch
    b = sample.batch
    dev = engine.DeviceBatch(b)
    mq = torch.from_numpy(sample.mq).cuda()
    dev.forward(mode=0)
    ref = dev.genotype(mq).clone()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.graph(g, stream=s):
        dev.forward(mode=0, stream=s)
        out = dev.genotype(mq, stream=s)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        dev.forward(mode=0)           # eager work between replays
        torch.cuda.synchronize()


def test_genotype_follows_forward_on_default_stream(sample):
    """Forward and genotype enqueued back to back on torch's default stream
    (handle NULL) with no synchronisation in between: the genotype kernels run
    after the forward has written every likelihood (the buffer starts as NaN,
    so reading it early would show)."""
    import torch
    dev = engine.DeviceBatch(sample.batch)
    mq = torch.from_numpy(sample.mq).cuda()
    dev.forward(mode=0)
    ref = dev.genotype(mq, stats=True)[0].clone()
    for _ in range(3):
        dev.out.fill_(float("nan"))
        dev.forward(mode=0)
        got = dev.genotype(mq)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)


def test_near_tie_guard_recomputes_in_fp64(sample):
    """Regions given a duplicate candidate haplotype have exactly tied
    genotypes: the guard flags them (with the rule of partial_anchor.py:119-133
    applied to the device's likelihoods) and replaces their outputs by the chain
    on all-FP64 likelihoods; every other region keeps its mixed-precision
    result."""
    import torch
    src = sample.batch
    regions, mqs = [], []
    for k in range(src.nregions):
        reads, quals, haps = src.region(k)
        if k % 5 == 0 and haps:
            haps = haps + [haps[0]]
        regions.append((reads, quals, haps))
        mqs.append(sample.mq[src.region_reads[k]:src.region_reads[k + 1]])
    b = workloads.from_regions(regions, src.gap, src.ext)
    mq_h = np.concatenate(mqs)
    mq = torch.from_numpy(mq_h).cuda()
    dev = engine.DeviceBatch(b)
    dev.forward(mode=0)
    plain, _, cert = dev.genotype(mq, values=True)
    plain = plain.clone()
    cert = cert.cpu().numpy()
    guarded, report = dev.genotype_guarded(mq, mode=0)
    g0 = plain.cpu().numpy()
    g1 = guarded.cpu().numpy()
    nh = np.diff(b.region_haps)
    off = engine.gl_offsets(nh)
    # the oracle's restatement of the rule on the device's own likelihoods and certainties
    flags, gaps = oracle.guard(b.region_reads, nh, off[:-1], g0, cert)
    np.testing.assert_array_equal(report["flags"].cpu().numpy(), flags)
    np.testing.assert_allclose(report["gaps"].cpu().numpy().reshape(-1, 2), gaps, rtol=1e-12, atol=1e-12)
    expected = list(np.flatnonzero(flags))
    flagged = [r["region"] for r in report["regions"]]
    assert flagged == expected
    assert flagged and set(flagged) & set(range(0, b.nregions, 5))   # the planted ties are found
    full = engine.DeviceBatch(b)
    full.forward(mode=0, precision="float64")
    g64 = full.genotype(mq).cpu().numpy()
    for k in range(b.nregions):
        sl = slice(off[k], off[k + 1])
        np.testing.assert_array_equal(g1[sl], g64[sl] if k in set(flagged) else g0[sl])


def test_near_tie_guard_allele_frequency_prior(sample):
    """The guard with partial_anchor.py:56-57's allele-frequency prior: flags
    and gaps equal the oracle's restatement on the same likelihoods, including
    regions where only the prior makes the posterior gap close (a
    heterozygote whose frequencies are set to cancel its likelihood lead)."""
    import torch
    b = sample.batch
    dev = engine.DeviceBatch(b)
    mq = torch.from_numpy(sample.mq).cuda()
    dev.forward(mode=0)
    rng = np.random.default_rng(17)
    af = rng.uniform(0, 0.6, b.nhaps)
    af[rng.random(b.nhaps) < 0.2] = 0.0          # unknown frequencies: floored at .001
    guarded, report = dev.genotype_guarded(mq, mode=0, allele_freq=af)
    gl, _, cert = dev.genotype(mq, values=True)
    nh = np.diff(b.region_haps)
    off = engine.gl_offsets(nh)
    flags, gaps = oracle.guard(b.region_reads, nh, off[:-1], gl.cpu().numpy(), cert.cpu().numpy(), af=af)
    np.testing.assert_array_equal(report["flags"].cpu().numpy(), flags)
    np.testing.assert_allclose(report["gaps"].cpu().numpy().reshape(-1, 2), gaps, rtol=1e-12, atol=1e-12)
    assert report["prior"] == "allele_frequency"
    uniform = oracle.guard(b.region_reads, nh, off[:-1], gl.cpu().numpy(), cert.cpu().numpy())[1]
    assert np.any(np.abs(gaps[:, 1] - uniform[:, 1]) > 1e-3)     # the prior changes the posterior gaps
    with pytest.raises(ValueError, match="one frequency per haplotype"):
        dev.genotype_guarded(mq, allele_freq=af[:-1])
