"""HIP engine vs the FP64 oracle and the reference's golden vectors (needs an MI355X).

Tolerances: FP32+FP64-rescue ("mixed") within 1e-4 log10 = 2.3026e-4 in ln
per pair (the engine's accuracy contract); all-FP64 within 1e-9 ln (the
reference FP64 CUDA path is held to 2e-9, test_accelerator.py:50)."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from tests.golden import fixtures as fx
from vardepth_amd import engine, workloads, _lib

pytestmark = pytest.mark.gpu
TOL_MIXED = 1e-4 * math.log(10)
TOL_FP64 = 1e-9
G = fx.load()


def close(actual, expected, tol):
    actual, expected = np.asarray(actual), np.asarray(expected)
    assert actual.shape == expected.shape
    inf_a, inf_e = np.isneginf(actual), np.isneginf(expected)
    np.testing.assert_array_equal(inf_a, inf_e)
    ok = ~inf_e
    assert np.isfinite(actual[ok]).all(), actual[ok][~np.isfinite(actual[ok])]
    err = np.abs(actual[ok] - expected[ok])
    assert err.size == 0 or err.max() <= tol, (err.max(), np.argmax(err))
    return float(err.max()) if err.size else 0.0


def region_batch_from_pairs(pairs, gap=0.001, ext=0.1):
    """Each pair as its own 1x1 region."""
    return workloads.from_regions([([s], [q], [h]) for s, q, h in pairs], gap, ext)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("precision,tol", [("mixed", TOL_MIXED), ("float64", TOL_FP64)])
def test_active_region_vs_oracle(mode, precision, tol):
    b = workloads.single_region(seed=41001)
    out, st = engine.forward(b, mode=mode, precision=precision, stats=True)
    reads, quals, haps = b.region(0)
    expect = np.array([oracle.dense(s, q, h, mode, b.gap, b.ext) for s, q in zip(reads, quals) for h in haps])
    err = close(out, expect, tol)
    print("mode %d %s: max |ln err| %.3g, flagged %d, fp64 %d" % (mode, precision, err, st["flagged"],
                                                                   st["fp64_pairs"]))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_golden_region_dense(mode):
    """The 64x8 region at g=.001, x=.1 against the reference's pairhmm_pairs outputs."""
    pairs = fx.pairs("region")
    R, H = (int(v) for v in G["region_shape"])
    b = workloads.from_regions([([pairs[i * H][0] for i in range(R)], [pairs[i * H][1] for i in range(R)],
                                 [pairs[j][2] for j in range(H)])])
    close(engine.forward(b, mode=mode, precision="float64"), G["region_dense_m%d" % mode], TOL_FP64)
    close(engine.forward(b, mode=mode, precision="mixed"), G["region_dense_m%d" % mode], TOL_MIXED)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_grid_shapes(mode):
    """Reference CUDA-test shapes (1,1)..(700,750), each pair its own region."""
    pairs = fx.pairs("grid")
    b = region_batch_from_pairs(pairs)
    close(engine.forward(b, mode=mode, precision="float64"), G["grid_dense_m%d" % mode], TOL_FP64)
    close(engine.forward(b, mode=mode, precision="mixed"), G["grid_dense_m%d" % mode], TOL_MIXED)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_edge_cases(mode):
    pairs = fx.pairs("edge")
    b = region_batch_from_pairs(pairs)
    close(engine.forward(b, mode=mode, precision="float64"), G["edge_dense_m%d" % mode], TOL_FP64)
    close(engine.forward(b, mode=mode, precision="mixed"), G["edge_dense_m%d" % mode], TOL_MIXED)


def test_precision_stress_rescue():
    """Low-quality 250 bp reads vs 1-2 kb indel-dense haplotypes: FP32 bulk, FP64 rescue."""
    b = workloads.precision_stress(seed=41003, regions=6)
    out, st = engine.forward(b, mode=0, precision="mixed", stats=True)
    expect, _ = oracle.regions_cpu(b, mode=0)
    err = close(out, expect, TOL_MIXED)
    assert st["flagged"] > 0 and st["fp64_pairs"] >= st["flagged"]
    print("stress: pairs %d flagged %d max err %.3g" % (st["pairs"], st["flagged"], err))


def test_fp64_routing_gives_fp64_values():
    """Classes routed straight to FP64 (vd_tuning.fwd_route64) return the
    all-FP64 values for their pairs, and the other classes their FP32 values;
    the autotuner keeps routing only for a batch that FP32 mostly flags."""
    from vardepth_amd import autotune
    base = autotune.get()
    b = workloads.uneven_regions(seed=5, regions=24)
    dev = engine.DeviceBatch(b)
    try:
        dev.forward(mode=0, precision="mixed")
        mixed = dev.result()
        dev.forward(mode=0, precision="float64")
        f64 = dev.result()
        rl = b.read_lengths()
        cls = (rl + 15) // 16  # rows per lane of each read (mode 0)
        routed = [r for r in range(1, 17) if r % 2 == 1]
        t = autotune.get()
        for r in routed:
            t.fwd_route64[r] = 1
        autotune.set_tuning(t)
        dev.forward(mode=0, precision="mixed")
        out = dev.result()
        oo = b.out_off()
        for k in range(b.nregions):
            r0, r1 = b.region_reads[k], b.region_reads[k + 1]
            nh = b.region_haps[k + 1] - b.region_haps[k]
            for i in range(r0, r1):
                sl = slice(oo[k] + (i - r0) * nh, oo[k] + (i - r0 + 1) * nh)
                want = f64[sl] if cls[i] in routed else mixed[sl]
                np.testing.assert_array_equal(out[sl], want)
        autotune.set_tuning(base)
        stress = engine.DeviceBatch(workloads.precision_stress(seed=41003, regions=8))
        rep = autotune.tune_forward(stress, mode=0, reps=1)
        assert rep["route64"]["flagged_frac"] > 0.25 and "routed_ms" in rep["route64"]
    finally:
        autotune.set_tuning(base)


def test_uneven_regions_subset():
    b = workloads.uneven_regions(seed=41002, regions=24)
    out, st = engine.forward(b, mode=0, precision="mixed", stats=True)
    expect, _ = oracle.regions_cpu(b, mode=0)
    err = close(out, expect, TOL_MIXED)
    print("uneven: pairs %d cells %d flagged %d max err %.3g ms %.2f" % (
        st["pairs"], st["cells"], st["flagged"], err, st["device_ms"]))


def test_output_indexing_is_exact():
    """Every output slot is written by exactly its own pair: permute regions and
    compare bit for bit (an FP64 result depends on its pair alone, not on the
    pair's unit mates or place in the work list)."""
    b = workloads.uneven_regions(seed=7, regions=12)
    out = engine.forward(b, mode=0, precision="float64")
    order = list(range(b.nregions))[::-1]
    b2 = b.subset(order)
    out2 = engine.forward(b2, mode=0, precision="float64")
    oo, oo2 = b.out_off(), b2.out_off()
    for new, k in enumerate(order):
        np.testing.assert_array_equal(out2[oo2[new]:oo2[new + 1]], out[oo[k]:oo[k + 1]])


def test_graph_capture_replays():
    """The whole forward captured in a HIP graph: every replay, with eager work
    and synchronisation in between, reproduces the eager result."""
    import torch
    b = workloads.uneven_regions(seed=5, regions=16)
    dev = engine.DeviceBatch(b)
    dev.forward(mode=0, precision="mixed")
    torch.cuda.synchronize()
    eager = dev.result()
    g = dev.graph(mode=0, precision="mixed")
    scratch = torch.zeros(1 << 20, device="cuda")
    for rep in range(4):
        dev.out.fill_(float("nan"))
        scratch.add_(1.0)  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dev.result(), eager)
    expect, _ = oracle.regions_cpu(b, mode=0)
    close(eager, expect, TOL_MIXED)


def test_plain_baseline_matches():
    """The one-pair-per-launch baseline computes the same likelihoods (FP32, no
    rescue: only pairs inside the FP32 window are compared -- the 64x8 region in
    semiglobal mode, and the short grid shapes in anchored mode, where 150 bp
    reads against longer haplotypes fall below it)."""
    pairs = fx.pairs("region")
    R, H = (int(v) for v in G["region_shape"])
    b = workloads.from_regions([([pairs[i * H][0] for i in range(R)], [pairs[i * H][1] for i in range(R)],
                                 [pairs[j][2] for j in range(H)])])
    out, ms = engine.forward_plain(b, mode=0, max_pairs=200)
    assert np.isnan(out[200:]).all() and ms > 0
    close(out[:200], G["region_dense_m0"][:200], TOL_MIXED)
    grid = fx.pairs("grid")
    short = [i for i, (s, q, h) in enumerate(grid) if 0 < len(s) <= 40]
    gb = region_batch_from_pairs([grid[i] for i in short])
    out1, _ = engine.forward_plain(gb, mode=1)
    close(out1, G["grid_dense_m1"][short], TOL_MIXED)


def test_class_overlap_keeps_results():
    """Consecutive class kernels on two streams (the default) give the same
    likelihoods as one stream, bit for bit outside the FP64-rescued pairs, in
    modes 0/1/2 with prefix sharing and the suffix continuation, with pair
    units, and from a captured graph."""
    import torch
    L = _lib.lib()
    prev = L.vd_class_overlap(-1)
    prev_share = L.vd_prefix_sharing(-1)
    try:
        b = workloads.uneven_regions(seed=12, regions=48)
        dev = engine.DeviceBatch(b)
        for share in (1, 0):
            L.vd_prefix_sharing(share)
            for mode in (0, 1, 2):
                L.vd_class_overlap(0)
                dev.forward(mode=mode)
                ref = dev.result()
                L.vd_class_overlap(1)
                st = dev.forward(mode=mode, stats=True)
                got = dev.result()
                np.testing.assert_array_equal(got, ref)   # NaN positions included
        L.vd_prefix_sharing(1)
        L.vd_class_overlap(0)
        dev.forward(mode=0)
        ref = dev.result()
        L.vd_class_overlap(1)
        g = dev.graph(mode=0)
        for _ in range(3):
            g.replay()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(dev.result(), ref)
            dev.forward(mode=0)   # eager work between replays
    finally:
        L.vd_class_overlap(prev)
        L.vd_prefix_sharing(prev_share)


def test_autotune_keeps_results():
    """Every launch choice gives the same likelihoods and gradients (with the
    suffix continuation off: it agrees to FP32 rounding, tested on its own)."""
    from vardepth_amd import autotune
    prev_sfx = _lib.lib().vd_suffix_sharing(0)
    prev_ov = _lib.lib().vd_class_overlap(-1)
    b = workloads.uneven_regions(seed=9, regions=24)
    dev = engine.DeviceBatch(b)
    base = autotune.get()
    dev.forward(mode=0)
    ref = dev.result()
    g_ref = dev.grad(mode=0).cpu().numpy()
    rep = autotune.tune_forward(dev, mode=0, reps=1)
    rep2 = autotune.tune_grad(dev, mode=0, reps=1)
    assert rep["choice"] and rep2["choice"]
    for cand in autotune.FWD_CANDIDATES:
        for chunk in (1, 8, 64):
            t = autotune.get()
            t.share_chunk = chunk
            for r in range(1, 17):
                t.fwd_unroll[r], t.fwd_blocks_per_cu[r], t.fwd_share[r] = cand
            autotune.set_tuning(t)
            dev.forward(mode=0)
            # FP32 sweeps agree bit for bit, and so do the pairs rescued in FP64
            np.testing.assert_array_equal(dev.result(), ref)
    for cand in autotune.GRAD_CANDIDATES:
        t = autotune.get()
        for r in range(1, 17):
            t.grad_variant[r], t.grad_blocks_per_cu[r] = cand
        autotune.set_tuning(t)
        g = dev.grad(mode=0).cpu().numpy()
        np.testing.assert_allclose(g, g_ref, rtol=1e-12, atol=1e-12)
    assert rep["class_overlap"] in (0, 1) and set(rep["overlap_ms"]) == {0, 1}
    autotune.set_tuning(base)
    _lib.lib().vd_suffix_sharing(prev_sfx)
    _lib.lib().vd_class_overlap(prev_ov)


def _sharing_cases():
    rng = np.random.default_rng(77)
    seq = lambda n: "".join(rng.choice(list("ACGT"), n))
    cases = [workloads.uneven_regions(seed=7, regions=48), workloads.chromosome_sample(mb=0.05, depth=30).batch]
    base = seq(300)
    reads = [base[i:i + L] for i, L in ((0, 40), (10, 150), (50, 250), (120, 100), (5, 17), (200, 96))]
    quals = [bytes(rng.integers(5, 41, len(r)).astype(np.uint8)) for r in reads]
    mut = lambda s, i, c: s[:i] + c + s[i + 1:]
    edge = [
        # prefix of another, duplicates, an empty haplotype, single-column divergence at the end
        (reads, quals, [base[:200], base, base, "", mut(base, 299, "A" if base[299] != "A" else "C"), base[:1]]),
        # many distinct forks: 40 haplotypes diverging at 40 different columns (beyond the 32 kept slots)
        (reads[:3], quals[:3], [mut(base, 5 * i + 3, "T" if base[5 * i + 3] != "T" else "G") for i in range(40)]),
        # more haplotypes than the planner ranks (swept without sharing)
        (reads[:1], quals[:1], [mut(base, i % 300, "ACGT"[i % 4]) for i in range(260)]),
        # identical haplotypes only
        (reads[:2], quals[:2], [base[:120]] * 5),
    ]
    cases.append(workloads.from_regions(edge))
    return cases


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_prefix_sharing_matches_full_sweeps_bit_for_bit(mode):
    L = _lib.lib()
    prev = L.vd_prefix_sharing(-1)
    prev_sfx = L.vd_suffix_sharing(0)
    try:
        for b in _sharing_cases():
            L.vd_prefix_sharing(0)
            full, st0 = engine.forward(b, mode=mode, stats=True)
            L.vd_prefix_sharing(1)
            shared, st1 = engine.forward(b, mode=mode, stats=True)
            assert st1["flagged"] == st0["flagged"]
            # the FP32 sweeps agree bit for bit, and so do the pairs rescued in FP64
            # (an FP64 result does not depend on the pair's unit mates or list order)
            np.testing.assert_array_equal(shared, full)
    finally:
        L.vd_prefix_sharing(prev)
        L.vd_suffix_sharing(prev_sfx)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_packed_pair_units_match_region_units_bit_for_bit(mode):
    """vd_tuning.fwd_pack: pair units whose 4 pairs come from different regions
    give the same likelihoods as region-bound pair units, bit for bit (FP64
    rescued pairs included), on uneven regions, one-pair regions (the
    true-haplotype batch), the precision-stress batch and the edge cases."""
    from vardepth_amd import autotune
    L = _lib.lib()
    prev = L.vd_prefix_sharing(-1)
    base = autotune.get()
    hmm = workloads.hmm_sampled_regions(seed=41004, regions=12, reads=30, haps=6)
    cases = _sharing_cases() + [workloads.true_pairs(hmm), workloads.precision_stress(seed=41003, regions=8)]
    try:
        L.vd_prefix_sharing(0)
        for b in cases:
            t = autotune.get()
            t.fwd_pack = 0
            autotune.set_tuning(t)
            ref, st0 = engine.forward(b, mode=mode, stats=True)
            t.fwd_pack = 1
            autotune.set_tuning(t)
            packed, st1 = engine.forward(b, mode=mode, stats=True)
            assert st1["flagged"] == st0["flagged"]
            assert st1["fp32_units"] <= st0["fp32_units"]
            np.testing.assert_array_equal(packed, ref)
    finally:
        autotune.set_tuning(base)
        L.vd_prefix_sharing(prev)


def test_many_regions_scan_in_several_chunks():
    """300,000 one-pair regions (5.4 M buckets, more tiles than one pass of the
    tile-sum scan covers): every pair lands in its own output slot with the
    oracle's value, with region-bound and with packed units."""
    from vardepth_amd import autotune
    rng = np.random.default_rng(300)
    n = 300_000
    rl = rng.integers(1, 41, n)
    hl = rng.integers(1, 61, n)
    bases = np.frombuffer(b"ACGT", np.uint8)
    reads = bases[rng.integers(0, 4, int(rl.sum()))]
    haps = bases[rng.integers(0, 4, int(hl.sum()))]
    ro = np.concatenate([[0], np.cumsum(rl)])
    ho = np.concatenate([[0], np.cumsum(hl)])
    b = workloads.RegionBatch(read_bytes=reads.copy(), qual_bytes=rng.integers(5, 41, int(rl.sum())).astype(np.uint8),
                              read_off=ro.astype(np.int64), hap_bytes=haps.copy(), hap_off=ho.astype(np.int64),
                              region_reads=np.arange(n + 1, dtype=np.int64),
                              region_haps=np.arange(n + 1, dtype=np.int64), gap=0.001, ext=0.1)
    assert b.nregions * 18 > 1024 * 4096
    sample = rng.choice(n, 400, replace=False)
    expect = np.array([oracle.dense(b.read_bytes[ro[i]:ro[i + 1]].tobytes().decode(),
                                    b.qual_bytes[ro[i]:ro[i + 1]].tobytes(),
                                    b.hap_bytes[ho[i]:ho[i + 1]].tobytes().decode(), 0, b.gap, b.ext)
                       for i in sample])
    base = autotune.get()
    try:
        outs = []
        for pack in (0, 1):
            t = autotune.get()
            t.fwd_pack = pack
            autotune.set_tuning(t)
            out = engine.forward(b, mode=0)
            close(out[sample], expect, TOL_MIXED)
            outs.append(out)
        np.testing.assert_array_equal(outs[0], outs[1])
    finally:
        autotune.set_tuning(base)


def _suffix_cases():
    rng = np.random.default_rng(91)
    seq = lambda n: "".join(rng.choice(list("ACGT"), n))
    tail = seq(120)
    left = seq(200)
    mut = lambda s, i, c: s[:i] + c + s[i + 1:]
    other = lambda s, i: mut(s, i, "A" if s[i] != "A" else "C")
    hap0 = left + tail
    reads = [hap0[i:i + L] for i, L in ((0, 60), (150, 150), (200, 120), (250, 70), (300, 20), (90, 255))]
    reads.append(seq(100))  # unrelated read: tiny likelihoods (rescued or clamped)
    quals = [bytes(rng.integers(5, 41, len(r)).astype(np.uint8)) for r in reads]
    edge = [
        # suffix of 120 after haplotypes that differ near the start, in the middle and just before it
        (reads, quals, [hap0, other(hap0, 3), other(hap0, 150), other(hap0, 199), left[:80] + tail]),
        # one haplotype is the suffix itself (it sweeps column 0 only)
        (reads[:3], quals[:3], [hap0, tail, other(hap0, 10)]),
        # identical haplotypes (the whole haplotype is the suffix)
        (reads[:4], quals[:4], [hap0] * 4),
        # common suffix of exactly 16 and of 15 (below the minimum: no continuation)
        (reads[:3], quals[:3], [seq(100) + tail[-16:], seq(90) + tail[-16:], seq(120) + tail[-16:]]),
        (reads[:3], quals[:3], [seq(100) + tail[-15:], "A" + seq(89) + tail[-15:]]),
        # many haplotypes sharing the suffix, beyond one read unit and beyond 32 forks
        (reads[:2], quals[:2], [other(hap0, 5 * i) for i in range(40)]),
    ]
    cases = [workloads.chromosome_sample(mb=0.05, depth=30).batch, workloads.uneven_regions(seed=8, regions=48),
             workloads.from_regions(edge)]
    return cases


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_suffix_continuation_matches_full_sweeps(mode):
    """The shared-suffix continuation (one adjoint sweep per read over the
    suffix every haplotype of a region ends in, then a dot product with each
    haplotype's state at m - S) agrees with full sweeps to FP32 rounding, and
    sweeps fewer cells where a suffix is shared."""
    from vardepth_amd import autotune
    L = _lib.lib()
    prev = L.vd_suffix_sharing(-1)
    base = autotune.get()
    try:
        t = autotune.get()
        # the suffix kernel for every read unit (2), with units of 4 and 2
        # haplotypes, and split with the prefix-only kernel by region (3)
        for share, chunk in ((2, 4), (2, 2), (3, 4)):
            t.share_chunk = chunk
            for r in range(1, 17):
                t.fwd_share[r] = share
            autotune.set_tuning(t)
            for i, b in enumerate(_suffix_cases()):
                L.vd_suffix_sharing(0)
                full, st0 = engine.forward(b, mode=mode, stats=True)
                L.vd_suffix_sharing(1)
                sfx, st1 = engine.forward(b, mode=mode, stats=True)
                assert st1["fp32_cells_swept"] <= st0["fp32_cells_swept"]
                if i != 1:  # the chromosome sample and the edge cases share suffixes
                    assert st1["fp32_cells_swept"] < st0["fp32_cells_swept"], (i, st0, st1)
                both = np.isfinite(full) & np.isfinite(sfx)
                assert np.array_equal(np.isfinite(full), np.isfinite(sfx))
                assert np.array_equal(np.isnan(full), np.isnan(sfx))
                np.testing.assert_allclose(sfx[both], full[both], rtol=0, atol=3e-5)
                # and against the oracle (the reference model in FP64)
                if i == 2:
                    expect = np.concatenate([
                        [oracle.dense(r, q, h, mode, b.gap, b.ext) for r, q in zip(*b.region(k)[:2])
                         for h in b.region(k)[2]] for k in range(b.nregions)])
                    close(sfx, expect, TOL_MIXED)
    finally:
        L.vd_suffix_sharing(prev)
        autotune.set_tuning(base)


def test_calls_on_different_streams_do_not_share_workspaces_concurrently():
    """Forwards of two different batches enqueued back to back on two streams
    with no synchronisation in between: the second call's stream waits for the
    first call's work (they share the engine's plan and sweep workspaces), so
    both results equal their eager values."""
    import torch
    a = engine.DeviceBatch(workloads.uneven_regions(seed=21, regions=96))
    b = engine.DeviceBatch(workloads.chromosome_sample(mb=0.05, depth=30).batch)
    a.forward(mode=0)
    ref_a = a.result().copy()
    b.forward(mode=0)
    ref_b = b.result().copy()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        a.out.fill_(float("nan"))
        b.out.fill_(float("nan"))
        torch.cuda.synchronize()
        a.forward(mode=0, stream=s1)
        b.forward(mode=0, stream=s2)
        a.forward(mode=0, stream=s2)   # and the first batch again, after the second
        torch.cuda.synchronize()
        # FP64-rescued pairs may move in the last bit with the rescue list's order
        np.testing.assert_array_equal(a.result(), ref_a)
        np.testing.assert_array_equal(b.result(), ref_b)


@pytest.mark.parametrize("mode", [0, 2])
def test_chromosome_sample_forward_vs_oracle(mode):
    """The forward over a 0.3 Mb slice of the chromosome-scale sample at 30x
    (prefix sharing and the shared-suffix continuation on, as by default)
    against the reference CPU chain restated in the oracle (trie, dense
    fallback), pair by pair; NaN and -inf positions must agree exactly."""
    b = workloads.chromosome_sample(mb=0.3, depth=30.0).batch
    out, st = engine.forward(b, mode=mode, precision="mixed", stats=True)
    expect, dense = oracle.regions_cpu(b, mode=mode)
    np.testing.assert_array_equal(np.isnan(out), np.isnan(expect))
    err = close(out, expect, TOL_MIXED)
    assert st["fp32_cells_swept"] < st["cells"]  # sharing took part
    print("chromosome 0.3 Mb mode %d: %d regions, %d pairs, max |ln err| %.3g, flagged %d, swept %.3g of %.3g" % (
        mode, b.nregions, st["pairs"], err, st["flagged"], st["fp32_cells_swept"], st["cells"]))


def test_graph_keeps_its_workspaces_and_parameters():
    """A captured forward replays its own parameters and buffers after a larger
    batch with other parameters grew and rewrote the engine's workspaces; a
    capture that would need other parameter tables (or larger buffers) than
    the last eager call left is refused with an error before it enqueues
    anything."""
    import torch
    small = workloads.uneven_regions(seed=5, regions=16)
    dev = engine.DeviceBatch(small)
    g = dev.graph(mode=0, precision="mixed")
    torch.cuda.synchronize()
    captured = dev.result()
    # a capture with other parameters than the last eager call would have to
    # upload tables inside the graph: refused
    big = engine.DeviceBatch(workloads.uneven_regions(seed=6, regions=96))
    g2 = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with pytest.raises(_lib.VdhmmError, match="stream capture"):
        with torch.cuda.graph(g2, stream=s):
            big.forward(mode=0, precision="mixed", stream=s, gap=0.02)
    del g2
    # an eager call of the larger batch with other parameters allocates its own workspaces and tables
    big.forward(mode=1, precision="mixed", gap=0.01, ext=0.3)
    torch.cuda.synchronize()
    assert big.max_hap >= dev.max_hap and big.npairs > 4 * dev.npairs
    for rep in range(3):
        dev.out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dev.result(), captured)
        big.forward(mode=0, precision="float64", gap=0.02, ext=0.2, err=np.full(64, 0.01))
        torch.cuda.synchronize()
    expect, _ = oracle.regions_cpu(small, mode=0)
    close(captured, expect, TOL_MIXED)


def test_haplotype_length_limit():
    """Batches with a haplotype longer than the kernels hold in LDS are refused
    with a clear error; one at the limit is scored."""
    lim = _lib.lib().vd_max_hap_len(0)
    assert 19000 < lim < 21000
    rng = np.random.default_rng(3)
    read = bytes(rng.choice(list(b"ACGT"), 100).astype(np.uint8))
    qual = bytes([30] * 100)
    ok = bytes(rng.choice(list(b"ACGT"), lim).astype(np.uint8))
    out = engine.forward(workloads.from_regions([([read], [qual], [ok])]), mode=0, precision="mixed")
    expect = oracle.dense(read, qual, ok, 0)
    assert abs(out[0] - expect) <= TOL_MIXED


@pytest.mark.parametrize("mode", [0, 1])
def test_reference_design_baseline_vs_oracle(mode):
    """The batched reference-GPU-design baseline computes the same model: its
    non-NaN pairs agree with the FP64 oracle to its FP32 accuracy, -inf for
    empty reads and haplotypes."""
    if mode == 0:
        b = workloads.uneven_regions(seed=3, regions=12)
    else:
        # anchored (global) alignment: haplotypes about as long as the reads,
        # otherwise the design's 2^-40 terminal rule sends most pairs to FP64
        rng = np.random.default_rng(3)
        regions = []
        for _ in range(8):
            L = int(rng.integers(80, 200))
            haps = ["".join(rng.choice(list("ACGT"), L)) for _ in range(3)]
            reads, quals = [], []
            for h in haps * 4:
                r = list(h)
                for i in rng.choice(L, 3, replace=False):
                    r[i] = "ACGT"[(int("ACGT".index(r[i])) + 1) % 4]
                reads.append("".join(r))
                quals.append(bytes(rng.integers(20, 41, L).astype(np.uint8)))
            regions.append((reads, quals, haps))
        b = workloads.from_regions(regions)
    edge = workloads.from_regions([(["ACGT", ""], [bytes([30] * 4), b""], ["ACGTA", ""])])
    for batch in (b, edge):
        out, ms, launches = engine.forward_refdesign(batch, mode=mode, budget_mb=64)
        expect, _ = oracle.regions_cpu(batch, mode=mode)
        assert launches >= 1 and ms > 0
        np.testing.assert_array_equal(np.isneginf(out), np.isneginf(expect))
        ok = np.isfinite(out)
        # anchored: each read against its own haplotype (a third of the pairs)
        # stays in the design's FP32 window; against unrelated ones it need not
        assert ok.sum() >= ((0.5 if mode == 0 else 1 / 3) * batch.npairs if batch is b else 1)
        assert np.abs(out[ok] - expect[ok]).max() <= TOL_MIXED
