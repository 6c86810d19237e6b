"""Adjoint kernels on the MI355X: gradients of sum w ln L against FP64 references.

References: the PyTorch float64 autograd restatement (oracle/torch_ref.py,
itself pinned to the reference's golden values) and central finite
differences of the C oracle with a relative step h = 1e-3.  Tolerance: 1e-3
relative to the largest component of each gradient group (d/dgap, d/dext,
d/derr), the accuracy contract of the gradient path; ln L within 1e-4 log10.
"""
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


def test_grad_vs_finite_differences():
    """Central differences (relative step 1e-3) of the C oracle's sum of ln L."""
    jobs = fx.pairs("region")[:48:3]
    gap, ext = 0.001, 0.1
    err = oracle.error_table()[:64].copy()
    out, g = engine.grad(batch_of(jobs, gap, ext), mode=0, err=err)

    def F(gp, xp, ep):
        e256 = np.array([ep[min(q, 63)] for q in range(256)])
        return sum(oracle.dense(s, q, h, 0, gp, xp, err=e256) for s, q, h in jobs)

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
