"""Adjoint kernels on the MI355X: gradients of sum w ln L against FP64 references.

References: the PyTorch float64 autograd restatement (oracle/torch_ref.py,
pinned by tests/test_oracle_torchref.py to the reference's golden values, to
exact rational derivatives of the brute-force path sum and to finite
differences) and central finite differences of the C oracle with a relative
step h = 1e-3.  Tolerance: 1e-3
relative to the largest component of each gradient group (d/dgap, d/dext,
d/derr), the accuracy contract of the gradient path; ln L within 1e-4 log10.
"""
import functools
import math

import numpy as np
import pytest

import oracle
from oracle import torch_ref
from tests.golden import fixtures as fx
from vardepth_amd import engine, workloads

pytestmark = pytest.mark.gpu
TOL_MIXED = 1e-4 * math.log(10)
RTOL = 1e-3


def rand_pairs(rng, count, nmax=40, mmax=50):
    out = []
    for _ in range(count):
        n = int(rng.integers(1, nmax + 1))
        m = int(rng.integers(1, mmax + 1))
        read = rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes()
        hap = rng.choice(list(b"ACGT"), m).astype(np.uint8).tobytes()
        if rng.random() < 0.7 and m > 4:  # mostly related pairs
            k = int(rng.integers(0, max(1, m - n + 1)))
            hap = (hap[:k] + read + hap[k:])[:m]
        qual = rng.integers(0, 64, n).astype(np.uint8).tobytes()
        out.append((read, qual, hap))
    return out


def batch_of(pairs, gap, ext):
    return workloads.from_regions([([s], [q], [h]) for s, q, h in pairs], gap, ext)


def check_grad(g, dg, dx, de, rtol=RTOL):
    assert abs(g[0] - dg) <= rtol * abs(dg) + 1e-9, (g[0], dg)
    assert abs(g[1] - dx) <= rtol * abs(dx) + 1e-9, (g[1], dx)
    scale = np.max(np.abs(de))
    err = np.max(np.abs(g[2:] - de))
    assert err <= rtol * scale + 1e-9, (err, scale, int(np.argmax(np.abs(g[2:] - de))))


def perturbed_err(seed):
    rng = np.random.default_rng(seed)
    return np.clip(oracle.error_table()[:64] * np.exp(rng.normal(0, 0.3, 64)), 1e-7, 0.7)


@pytest.mark.parametrize("mode", [0, 1])
def test_grad_matches_autograd_small(mode):
    rng = np.random.default_rng(100 + mode)
    pairs = rand_pairs(rng, 40)
    err = perturbed_err(1)
    gap, ext = 0.01, 0.2
    w = rng.uniform(0.5, 2.0, len(pairs))
    out, g, st = engine.grad(batch_of(pairs, gap, ext), mode=mode, err=err, weights=w, stats=True)
    v, dg, dx, de = torch_ref.value_and_grad(pairs, mode, gap, ext, err, w)
    np.testing.assert_allclose(out, v, rtol=0, atol=TOL_MIXED)
    check_grad(g, dg, dx, de)
    print("mode %d: dg %.6g/%.6g dx %.6g/%.6g |de| %.3g flagged %d" % (mode, g[0], dg, g[1], dx,
                                                                      np.max(np.abs(de)), st["flagged"]))


@pytest.mark.parametrize("gap,ext", [(1e-4, 1e-3), (0.001, 0.02), (0.05, 0.9)])
@pytest.mark.parametrize("mode", [0, 1])
def test_grad_extreme_transitions(mode, gap, ext):
    """Gap and extension at the corners of their range.  A small extension
    makes the mu part of the insertion sums (mu = x/4 over the normaliser)
    small against their beta part, which the kernel forms as a difference of
    two sums; a large one makes it dominant."""
    rng = np.random.default_rng(300 + mode)
    pairs = rand_pairs(rng, 24)
    err = perturbed_err(2)
    w = rng.uniform(0.5, 2.0, len(pairs))
    out, g = engine.grad(batch_of(pairs, gap, ext), mode=mode, err=err, weights=w)
    v, dg, dx, de = torch_ref.value_and_grad(pairs, mode, gap, ext, err, w)
    np.testing.assert_allclose(out, v, rtol=0, atol=TOL_MIXED)
    check_grad(g, dg, dx, de)


@pytest.mark.parametrize("mode", [0, 1])
def test_grad_region_vs_autograd(mode):
    """8 reads x 8 haplotypes of the golden 64x8 region (150 bp reads, 200-400 bp haplotypes)."""
    jobs = fx.pairs("region")
    R, H = (int(v) for v in G_SHAPE())
    reads = [jobs[i * H][0] for i in range(0, R, 8)]
    quals = [jobs[i * H][1] for i in range(0, R, 8)]
    haps = [jobs[j][2] for j in range(H)]
    b = workloads.from_regions([(reads, quals, haps)], 0.001, 0.1)
    err = oracle.error_table()[:64]
    out, g = engine.grad(b, mode=mode, err=err)
    pairs = [(s, q, h) for s, q in zip(reads, quals) for h in haps]
    v, dg, dx, de = torch_ref.value_and_grad(pairs, mode, 0.001, 0.1, err)
    np.testing.assert_allclose(out, v, rtol=0, atol=TOL_MIXED)
    check_grad(g, dg, dx, de)


def G_SHAPE():
    return fx.load()["region_shape"]


@pytest.mark.parametrize("mode", [0, 1])
def test_grad_vs_finite_differences(mode):
    """Central differences (relative step 1e-3) of the C oracle's sum of ln L."""
    jobs = fx.pairs("region")[:48:3]
    gap, ext = 0.001, 0.1
    err = oracle.error_table()[:64].copy()
    out, g = engine.grad(batch_of(jobs, gap, ext), mode=mode, err=err)

    def F(gp, xp, ep):
        e256 = np.array([ep[min(q, 63)] for q in range(256)])
        return sum(oracle.dense(s, q, h, mode, gp, xp, err=e256) for s, q, h in jobs)

    h = 1e-3
    fd_g = (F(gap * (1 + h), ext, err) - F(gap * (1 - h), ext, err)) / (2 * h * gap)
    fd_x = (F(gap, ext * (1 + h), err) - F(gap, ext * (1 - h), err)) / (2 * h * ext)
    assert abs(g[0] - fd_g) <= RTOL * abs(fd_g), (g[0], fd_g)
    assert abs(g[1] - fd_x) <= RTOL * abs(fd_x), (g[1], fd_x)
    used = sorted({min(c, 63) for _, q, _ in jobs for c in q})
    fd_e = {}
    for qb in used:
        ep, em = err.copy(), err.copy()
        ep[qb] *= 1 + h
        em[qb] *= 1 - h
        fd_e[qb] = (F(gap, ext, ep) - F(gap, ext, em)) / (2 * h * err[qb])
    scale = max(abs(v) for v in fd_e.values())
    for qb, v in fd_e.items():
        assert abs(g[2 + qb] - v) <= RTOL * scale, (qb, g[2 + qb], v)
    unused = [q for q in range(64) if q not in used]
    assert np.all(g[2 + np.array(unused)] == 0)


def test_grad_linear_in_weights():
    rng = np.random.default_rng(7)
    pairs = rand_pairs(rng, 30)
    b = batch_of(pairs, 0.003, 0.15)
    w1, w2 = rng.uniform(-1, 1, len(pairs)), rng.uniform(-1, 1, len(pairs))
    _, g1 = engine.grad(b, weights=w1)
    _, g2 = engine.grad(b, weights=w2)
    _, g12 = engine.grad(b, weights=w1 + w2)
    scale = np.max(np.abs(g12))
    assert np.max(np.abs(g1 + g2 - g12)) <= 1e-5 * scale
    _, g0 = engine.grad(b, weights=np.zeros(len(pairs)))
    assert np.all(g0 == 0)


def test_grad_fp64_rescue():
    """Low-quality 250 bp reads against 1-2 kb haplotypes leave the FP32 window;
    their gradients come from the FP64 adjoint kernel."""
    rng = np.random.default_rng(5)
    pairs = []
    for k in range(6):
        read = rng.choice(list(b"ACGT"), 250).astype(np.uint8).tobytes()
        hap = rng.choice(list(b"ACGT"), 300 + 40 * k).astype(np.uint8).tobytes()
        if k % 2:  # related pairs stay in the FP32 window
            hap = hap[:20] + read + hap[270:]
        pairs.append((read, rng.integers(2, 6, 250).astype(np.uint8).tobytes(), hap))
    b = batch_of(pairs, 0.001, 0.1)
    out, g, st = engine.grad(b, mode=0, stats=True)
    assert st["flagged"] >= 3 and st["fp64_bad"] == 0
    v, dg, dx, de = torch_ref.value_and_grad(pairs, 0, b.gap, b.ext, oracle.error_table()[:64])
    np.testing.assert_allclose(out, v, rtol=0, atol=TOL_MIXED)
    check_grad(g, dg, dx, de)


def long_pairs(rng, lengths, extra=(0, 300)):
    """Reads beyond 256 rows against haplotypes `extra` bases longer, mostly
    placed in their haplotype with a few substitutions, a few against
    unrelated haplotypes."""
    out = []
    for k, n in enumerate(lengths):
        m = n + int(rng.integers(*extra))
        read = rng.choice(list(b"ACGT"), n).astype(np.uint8)
        if k % 4 != 3 and m > n:
            hap = rng.choice(list(b"ACGT"), m).astype(np.uint8)
            o = int(rng.integers(0, m - n + 1))
            hap[o:o + n] = read
            for i in rng.choice(n, 4, replace=False):
                read[i] = b"ACGT"[(b"ACGT".index(int(read[i])) + 1) % 4]
        else:
            hap = rng.choice(list(b"ACGT"), m).astype(np.uint8)
        out.append((read.tobytes(), rng.integers(10, 41, n).astype(np.uint8).tobytes(), hap.tobytes()))
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_grad_long_reads_tiled(mode):
    """Reads of 257-700 bases (up to 11 tiles of 64 rows, unit mates with
    different tile counts) run in the tiled FP64 adjoint, together with short
    reads in the FP32 kernels: ln L and gradients against the float64
    autograd restatement."""
    rng = np.random.default_rng(40 + mode)
    lengths = [257, 300, 700, 380, 450, 640, 511, 600, 256 - mode, 320]
    # anchored: haplotypes only slightly longer than the reads, so that the
    # terminal stays above the reference's 1e-300 clamp the restatement omits
    pairs = long_pairs(rng, lengths, (0, 300) if mode == 0 else (1, 24))
    pairs += rand_pairs(rng, 6)
    err = perturbed_err(3)
    gap, ext = 0.004, 0.15
    w = rng.uniform(0.5, 2.0, len(pairs))
    long_ = np.array([len(p_[0]) + mode > 256 for p_ in pairs])
    only, wl = [p_ for p_, k in zip(pairs, long_) if k], w[long_]
    short, ws = [p_ for p_, k in zip(pairs, long_) if not k], w[~long_]
    vl, dgl, dxl, del_ = torch_ref.value_and_grad(only, mode, gap, ext, err, wl)
    vs, dgs, dxs, des = torch_ref.value_and_grad(short, mode, gap, ext, err, ws)
    # the long pairs alone: all FP64, ln L and the gradient to FP64 rounding
    outl, gl, st = engine.grad(batch_of(only, gap, ext), mode=mode, err=err, weights=wl, stats=True)
    np.testing.assert_allclose(outl, vl, rtol=0, atol=1e-9)
    check_grad(gl, dgl, dxl, del_, rtol=1e-7)
    assert st["fp64_bad"] == 0
    # with the short reads (FP32 kernels) in one call: the sum of both
    out, g = engine.grad(batch_of(pairs, gap, ext), mode=mode, err=err, weights=w)
    np.testing.assert_allclose(out[long_], vl, rtol=0, atol=1e-9)
    np.testing.assert_allclose(out[~long_], vs, rtol=0, atol=TOL_MIXED)
    check_grad(g, dgl + dgs, dxl + dxs, del_ + des)
    # a gradient is a sum over pairs: each long pair alone
    parts = [engine.grad(batch_of([p_], gap, ext), mode=mode, err=err, weights=wl[i:i + 1])[1]
             for i, p_ in enumerate(only)]
    np.testing.assert_allclose(gl, np.sum(parts, axis=0), rtol=1e-9, atol=1e-9 * np.max(np.abs(gl)))


def test_grad_empty_inputs():
    rng = np.random.default_rng(3)
    pairs = rand_pairs(rng, 6)
    regions = [([s], [q], [h]) for s, q, h in pairs] + [([b""], [b""], [b"ACGT"]), ([b"ACG"], [b"\x1e" * 3], [b""])]
    b = workloads.from_regions(regions, 0.001, 0.1)
    out, g = engine.grad(b)
    assert np.isneginf(out[-2:]).all()
    _, g_ref = engine.grad(batch_of(pairs, 0.001, 0.1))
    np.testing.assert_allclose(g, g_ref, rtol=1e-6, atol=1e-9)


def test_autograd_function():
    import torch
    from vardepth_amd import autograd
    rng = np.random.default_rng(11)
    pairs = rand_pairs(rng, 24)
    b = batch_of(pairs, 0.01, 0.2)
    dev = engine.DeviceBatch(b)
    err = perturbed_err(2)
    gap = torch.tensor(0.01, dtype=torch.float32, requires_grad=True)
    ext = torch.tensor(0.2, dtype=torch.float32, requires_grad=True)
    e = torch.tensor(err, dtype=torch.float32, requires_grad=True)
    ll = autograd.loglik(gap, ext, e, dev, 0)
    w = torch.tensor(rng.uniform(0.5, 2.0, len(pairs)), dtype=torch.float64, device=ll.device)
    loss = -(w * ll).sum()
    loss.backward()
    v, dg, dx, de = torch_ref.value_and_grad(pairs, 0, gap.item(), ext.item(), e.detach().double().numpy(),
                                             -w.cpu().numpy())
    np.testing.assert_allclose(ll.detach().cpu().numpy(), v, rtol=0, atol=TOL_MIXED)
    check_grad(np.concatenate([[gap.grad.item(), ext.grad.item()], e.grad.numpy()]), dg, dx, de)
    # the parameter module chains the sigmoid reparametrisation through torch
    mod = autograd.HmmParameters(0.01, 0.2, err)
    (-(mod(dev, 0)).sum()).backward()
    g_, x_, e_ = mod.values()
    assert mod.a.grad is not None and mod.c.grad.shape == (64,)
    assert abs(float(g_) - 0.01) < 1e-6 and abs(float(x_) - 0.2) < 1e-6


def test_autograd_fused_matches_two_calls():
    """loglik(fused=True): the forward's unit-weight adjoint, scaled, gives the
    gradient of a uniform loss (-mean ln L) as the forward + weighted adjoint
    does; a non-uniform incoming gradient falls back to the weighted adjoint."""
    import torch
    from vardepth_amd import autograd
    rng = np.random.default_rng(13)
    pairs = rand_pairs(rng, 24)
    dev = engine.DeviceBatch(batch_of(pairs, 0.01, 0.2))
    err = perturbed_err(3)
    w = torch.tensor(rng.uniform(0.5, 2.0, len(pairs)), dtype=torch.float64)

    def grads(fused, weighted):
        gap = torch.tensor(0.01, dtype=torch.float32, requires_grad=True)
        ext = torch.tensor(0.2, dtype=torch.float32, requires_grad=True)
        e = torch.tensor(err, dtype=torch.float32, requires_grad=True)
        ll = autograd.loglik(gap, ext, e, dev, 0, fused=fused)
        loss = -(w.to(ll.device) * ll).sum() if weighted else -ll.sum() / len(pairs)
        loss.backward()
        return ll.detach().cpu().numpy(), np.concatenate([[gap.grad.item(), ext.grad.item()], e.grad.numpy()])
    for weighted in (False, True):
        ll0, g0 = grads(False, weighted)
        ll1, g1 = grads(True, weighted)
        np.testing.assert_allclose(ll1, ll0, rtol=0, atol=TOL_MIXED)
        np.testing.assert_allclose(g1, g0, rtol=1e-5, atol=1e-7 * np.max(np.abs(g0)))


def test_adam_recovers_parameters():
    """200 Adam steps through the HIP forward/adjoint from a perturbed start move
    gap, continuation and the populated error bins towards the generator's."""
    from vardepth_amd import fit
    true_gap, true_ext = 0.01, 0.3
    err_true = np.power(10.0, -np.minimum(np.arange(64), 60) / 10.0)
    b = workloads.hmm_sampled_regions(seed=41004, regions=48, sample_gap=true_gap, sample_ext=true_ext)
    dev = engine.DeviceBatch(workloads.true_pairs(b))
    gap0, ext0, err0 = fit.perturbed_start(41004, true_gap, true_ext, err_true)
    _, hist = fit.fit(dev, gap0, ext0, err0, steps=200, lr=0.05)
    first, last = hist[0], hist[-1]
    assert last["loss"] < first["loss"]
    assert abs(last["gap"] - true_gap) < 0.3 * abs(first["gap"] - true_gap), (first["gap"], last["gap"])
    assert abs(last["ext"] - true_ext) < 0.3 * abs(first["ext"] - true_ext), (first["ext"], last["ext"])
    used = np.arange(10, 41)
    le = lambda e: np.mean(np.abs(np.log10(np.asarray(e)[used]) - np.log10(err_true[used])))
    assert le(last["err"]) < 0.5 * le(first["err"]), (le(first["err"]), le(last["err"]))
    print("gap %.4g -> %.4g, ext %.3g -> %.3g, log10 err %.3f -> %.3f" % (
        first["gap"], last["gap"], first["ext"], last["ext"], le(first["err"]), le(last["err"])))


# --- the differentiation workload: 1024 HMM-sampled regions, 2.05 M pairs ---------------------

@functools.lru_cache(maxsize=1)
def hmm_regions_1024():
    return workloads.hmm_sampled_regions(seed=41004, regions=1024)


def _slice(b, regions, reads, haps):
    picked = []
    for k in regions:
        rs, qs, hs = b.region(k)
        picked.append((rs[:reads], qs[:reads], hs[:haps]))
    jobs = [(s.encode(), q, h.encode()) for rs, qs, hs in picked for s, q in zip(rs, qs) for h in hs]
    return workloads.from_regions(picked, b.gap, b.ext), jobs


@pytest.mark.parametrize("mode", [0, 1])
def test_hmm_regions_slice_vs_autograd_and_fd(mode):
    """Reads of three of the 1024 regions (first, middle, last) against their
    first haplotypes: gradients against the float64 autograd restatement and
    against central differences of the C oracle."""
    b = hmm_regions_1024()
    sl, jobs = _slice(b, (0, 511, 1023), 4, 3)
    err = oracle.error_table()[:64].copy()
    out, g, st = engine.grad(sl, mode=mode, err=err, stats=True)
    v, dg, dx, de = torch_ref.value_and_grad(jobs, mode, b.gap, b.ext, err)
    np.testing.assert_allclose(out, v, rtol=0, atol=TOL_MIXED)
    check_grad(g, dg, dx, de)

    def F(gp, xp, ep):
        e256 = np.array([ep[min(q, 63)] for q in range(256)])
        return sum(oracle.dense(s, q, h, mode, gp, xp, err=e256) for s, q, h in jobs)

    h = 1e-3
    fd_g = (F(b.gap * (1 + h), b.ext, err) - F(b.gap * (1 - h), b.ext, err)) / (2 * h * b.gap)
    fd_x = (F(b.gap, b.ext * (1 + h), err) - F(b.gap, b.ext * (1 - h), err)) / (2 * h * b.ext)
    assert abs(g[0] - fd_g) <= RTOL * abs(fd_g), (g[0], fd_g)
    assert abs(g[1] - fd_x) <= RTOL * abs(fd_x), (g[1], fd_x)
    for qb in (10, 25, 40):
        ep, em = err.copy(), err.copy()
        ep[qb] *= 1 + h
        em[qb] *= 1 - h
        fd = (F(b.gap, b.ext, ep) - F(b.gap, b.ext, em)) / (2 * h * err[qb])
        assert abs(g[2 + qb] - fd) <= RTOL * np.max(np.abs(de)), (qb, g[2 + qb], fd)
    print("mode %d: %d pairs, %d flagged, dg %.6g dx %.6g" % (mode, sl.npairs, st["flagged"], g[0], g[1]))


def test_hmm_regions_full_batch_truth_weights():
    """The gradient over all 2.05 M pairs with weight 1 on each read's true
    haplotype and 0 elsewhere equals the gradient of the true-pair batch, and
    the ln L of the true pairs agree between the two layouts."""
    b = hmm_regions_1024()
    tp = workloads.true_pairs(b)
    oo = b.out_off()
    w = np.zeros(b.npairs)
    idx = []
    for k in range(b.nregions):
        r0, r1, h0 = b.region_reads[k], b.region_reads[k + 1], b.region_haps[k]
        H = b.region_haps[k + 1] - h0
        local = b.read_true_hap[r0:r1] - h0
        idx.append(oo[k] + np.arange(r1 - r0) * H + local)
    idx = np.concatenate(idx)
    w[idx] = 1.0
    out_full, g_full, st = engine.grad(b, weights=w, stats=True)
    out_true, g_true = engine.grad(tp)
    assert st["pairs"] == 2048000 and st["fp64_bad"] == 0
    assert np.isfinite(out_full).all()
    # true pairs, in the true-pair batch's order (region, haplotype, read)
    order = []
    for k in range(b.nregions):
        r0, r1, h0 = b.region_reads[k], b.region_reads[k + 1], b.region_haps[k]
        H = b.region_haps[k + 1] - h0
        local = b.read_true_hap[r0:r1] - h0
        for h in range(H):
            sel = np.flatnonzero(local == h)
            order.append(oo[k] + sel * H + h)
    np.testing.assert_allclose(out_full[np.concatenate(order)], out_true, rtol=0, atol=TOL_MIXED)
    scale = np.max(np.abs(g_true[2:]))
    assert abs(g_full[0] - g_true[0]) <= 1e-4 * abs(g_true[0])
    assert abs(g_full[1] - g_true[1]) <= 1e-4 * abs(g_true[1])
    assert np.max(np.abs(g_full[2:] - g_true[2:])) <= 1e-4 * scale


def test_adam_recovers_parameters_1024_regions():
    """200 Adam steps on the true-haplotype loss of all 1024 regions
    (102,400 reads) from the perturbed start."""
    from vardepth_amd import fit
    true_gap, true_ext = 0.01, 0.3
    err_true = np.power(10.0, -np.minimum(np.arange(64), 60) / 10.0)
    dev = engine.DeviceBatch(workloads.true_pairs(hmm_regions_1024()))
    gap0, ext0, err0 = fit.perturbed_start(41004, true_gap, true_ext, err_true)
    _, hist = fit.fit(dev, gap0, ext0, err0, steps=200, lr=0.05)
    first, last = hist[0], hist[-1]
    assert last["loss"] < first["loss"]
    assert abs(last["gap"] - true_gap) < 0.3 * abs(first["gap"] - true_gap), (first["gap"], last["gap"])
    assert abs(last["ext"] - true_ext) < 0.3 * abs(first["ext"] - true_ext), (first["ext"], last["ext"])
    used = np.arange(10, 41)
    le = lambda e: np.mean(np.abs(np.log10(np.asarray(e)[used]) - np.log10(err_true[used])))
    assert le(last["err"]) < 0.5 * le(first["err"]), (le(first["err"]), le(last["err"]))
    print("1024 regions: gap %.4g -> %.4g, ext %.3g -> %.3g, log10 err %.3f -> %.3f" % (
        first["gap"], last["gap"], first["ext"], last["ext"], le(first["err"]), le(last["err"])))
