"""The drop-in entry points on the MI355X against the reference's golden vectors.

C-ABI symbols (called through the reference's own ctypes prototypes, as
bound by vardepth_amd.native) and the Python mirrors of accelerator.py,
cpu_batch.py, trie_vector.py, whole_pairhmm.py and base/caller.py.

Tolerances: FP64 device paths within 1e-9 in ln (the reference holds its FP64
CUDA path to 2e-9, test_accelerator.py:50); FP32 + FP64-rescue paths within
1e-4 log10 = 2.3026e-4 in ln.  -inf and NaN positions must match exactly."""
import ctypes
import math

import numpy as np
import pytest

import oracle

from tests.golden import fixtures as fx
from vardepth_amd import accelerator, batch, caller, native, trie_vector, whole_pairhmm

pytestmark = pytest.mark.gpu
G = fx.load()
TOL64 = 1e-9
TOL_MIXED = 1e-4 * math.log(10)
P = ctypes.POINTER


def close(actual, expected, tol):
    actual, expected = np.asarray(actual, np.float64), np.asarray(expected, np.float64)
    assert actual.shape == expected.shape
    np.testing.assert_array_equal(np.isnan(actual), np.isnan(expected))
    np.testing.assert_array_equal(np.isneginf(actual), np.isneginf(expected))
    ok = np.isfinite(expected)
    assert np.isfinite(actual[ok]).all()
    err = np.abs(actual[ok] - expected[ok])
    assert err.size == 0 or err.max() <= tol, (err.max(), int(np.argmax(err)))


def c_pairs(jobs, mode, lanes=0):
    n = len(jobs)
    reads = (ctypes.c_char_p * n)(*[j[0] for j in jobs])
    quals = (ctypes.c_char_p * n)(*[j[1] for j in jobs])
    haps = (ctypes.c_char_p * n)(*[j[2] for j in jobs])
    rn = (ctypes.c_int * n)(*[len(j[0]) for j in jobs])
    hn = (ctypes.c_int * n)(*[len(j[2]) for j in jobs])
    out = np.empty(n, np.float64)
    code = native.cpu_batch_library().pairhmm_pairs(reads, ctypes.cast(quals, P(P(ctypes.c_ubyte))), rn, haps, hn, n,
                                                    mode, lanes, out.ctypes.data_as(P(ctypes.c_double)))
    assert code == 0
    return out


def c_trie(reads, quals, haps, mode, lanes=0):
    nr, nh = len(reads), len(haps)
    rp = (ctypes.c_char_p * max(nr, 1))(*reads)
    qp = (ctypes.c_char_p * max(nr, 1))(*quals)
    hp = (ctypes.c_char_p * max(nh, 1))(*haps)
    rn = (ctypes.c_int * max(nr, 1))(*[len(r) for r in reads])
    hn = (ctypes.c_int * max(nh, 1))(*[len(h) for h in haps])
    out = np.empty((nr, nh), np.float64)
    cells = ctypes.c_uint64()
    code = native.trie_library().trie_many(rp, ctypes.cast(qp, P(P(ctypes.c_ubyte))), rn, nr, hp, hn, nh, mode,
                                           lanes, out.ctypes.data_as(P(ctypes.c_double)), ctypes.byref(cells))
    assert code == 0
    return out, cells.value


def scalar(lib, name, jobs, gap=0.001, ext=0.1):
    fn = getattr(lib, name)
    return np.array([fn(s, ctypes.cast(ctypes.c_char_p(q), P(ctypes.c_ubyte)), len(s), h, len(h), gap, ext)
                     for s, q, h in jobs])


# --- pairhmm_pairs (batch_hmm.cpp:192) ------------------------------------------------------------

@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("prefix", ["grid", "edge", "region", "stress", "reuse"])
def test_pairhmm_pairs(prefix, mode):
    close(c_pairs(fx.pairs(prefix), mode), G["%s_dense_m%d" % (prefix, mode)], TOL64)


def test_pairhmm_pairs_lanes_do_not_change_values():
    jobs = fx.pairs("grid")
    for lanes in (2, 4, 8):
        close(c_pairs(jobs, 0, lanes), G["grid_dense_m0"], TOL64)


# --- trie_many (trie_vector.cpp:284) ----------------------------------------------------------------

@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("shape", range(6))
def test_trie_many_grid(shape, mode):
    jobs = fx.pairs("grid")[6 * shape:6 * shape + 6]
    out, cells = c_trie([j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs], mode)
    close(out, G["grid_trie%d_m%d" % (shape, mode)], TOL_MIXED)
    assert cells == int(G["grid_trie%d_m%d_cells" % (shape, mode)][0])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_trie_many_edge_and_region(mode):
    reads = [j for j in fx.pairs("edge") if j[0]]
    out, cells = c_trie([j[0] for j in reads], [j[1] for j in reads], fx.strings("edge_trie_haps"), mode)
    close(out, G["edge_trie_m%d" % mode], TOL_MIXED)
    assert cells == int(G["edge_trie_m%d_cells" % mode][0])
    jobs = fx.pairs("region")
    R, H = (int(v) for v in G["region_shape"])
    out, cells = c_trie([jobs[i * H][0] for i in range(R)], [jobs[i * H][1] for i in range(R)],
                        [jobs[j][2] for j in range(H)], mode)
    close(out, G["region_trie_m%d" % mode], TOL_MIXED)
    assert cells == int(G["region_trie_m%d_cells" % mode][0])


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("lanes", [0, 2, 4])
def test_trie_many_shared_suffix(mode, lanes):
    jobs = fx.pairs("suffix")
    out, cells = c_trie([j[0] for j in jobs], [j[1] for j in jobs], fx.strings("suffix_haps"), mode, lanes)
    close(out, G["suffix_trie_m%d_l%d" % (mode, lanes)], TOL_MIXED)
    assert cells == int(G["suffix_trie_m%d_l%d_cells" % (mode, lanes)][0])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_trie_many_underflow_is_nan(mode):
    vals = [c_trie([s], [q], [h], mode)[0][0, 0] for s, q, h in fx.pairs("reuse")]
    close(vals, G["reuse_trie_m%d" % mode], TOL_MIXED)


def test_trie_many_float64_env(monkeypatch):
    monkeypatch.setenv("VARDEPTH_CUDA_PRECISION", "float64")
    jobs = fx.pairs("region")
    R, H = (int(v) for v in G["region_shape"])
    out, _ = c_trie([jobs[i * H][0] for i in range(R)], [jobs[i * H][1] for i in range(R)],
                    [jobs[j][2] for j in range(H)], 0)
    close(out, G["region_trie_m0"], 1e-8)  # trie vs dense FP64 (test_cpu_reuse.py:21 holds 1e-9 on short cases)


# --- parametric scalar / batch entry points (wide.cpp, core.cpp, anchored.cpp) --------------------------

@pytest.mark.parametrize("prefix", ["grid", "edge"])
def test_scalar_entry_points(prefix):
    jobs = fx.pairs(prefix)
    close(scalar(native.wide_library(), "read_haplotype", jobs), G[prefix + "_wide"], TOL64)
    close(scalar(native.kernel(), "anchored_pairhmm", jobs), G[prefix + "_anch"], TOL64)
    close(scalar(native.core(), "read_haplotype", jobs), G[prefix + "_core"], TOL64)


def test_parametric_gap_ext():
    jobs = fx.pairs("param")
    for gi, g in enumerate(G["param_gaps"]):
        for xi, x in enumerate(G["param_exts"]):
            close(scalar(native.wide_library(), "read_haplotype", jobs, g, x), G["param_wide_g%d_x%d" % (gi, xi)], TOL64)
            close(scalar(native.core(), "read_haplotype", jobs, g, x), G["param_core_g%d_x%d" % (gi, xi)], TOL64)
            close(scalar(native.kernel(), "anchored_pairhmm", jobs, g, x), G["param_anchored_g%d_x%d" % (gi, xi)],
                  TOL64)


def test_batch_entry_points_region_q45():
    """One read x 8 haplotypes per call at g=10^-4.5, x=0.1 (the 64x8 region)."""
    jobs = fx.pairs("region")
    R, H = (int(v) for v in G["region_shape"])
    meta = fx.meta()["sets"]["region"]
    gap, ext = meta["gap"], meta["ext"]
    haps = [jobs[j][2] for j in range(H)]
    hp = (ctypes.c_char_p * H)(*haps)
    hn = (ctypes.c_int * H)(*[len(h) for h in haps])
    wide, core = native.wide_library(), native.core()
    got_w, got_wm, got_c = [], [], []
    for i in range(R):
        s, q = jobs[i * H][0], jobs[i * H][1]
        qp = ctypes.cast(ctypes.c_char_p(q), P(ctypes.c_ubyte))
        for fn, dst, extra in ((wide.read_haplotype_batch, got_w, ()), (wide.read_haplotype_batch_mode, got_wm, (8,)),
                               (core.read_haplotype_batch, got_c, ())):
            out = np.empty(H)
            fn(s, qp, len(s), hp, hn, H, gap, ext, out.ctypes.data_as(P(ctypes.c_double)), *extra)
            dst.append(out)
    close(np.concatenate(got_w), G["region_wide_q45"], TOL64)
    close(np.concatenate(got_wm), G["region_wide_q45"], TOL64)
    close(np.concatenate(got_c), G["region_core_q45"], TOL64)
    assert wide.wide_capability() == 8


# --- Python mirrors -------------------------------------------------------------------------------

def _str_jobs(prefix):
    return [(s.decode(), q, h.decode()) for s, q, h in fx.pairs(prefix)]


@pytest.mark.parametrize("precision,tol", [("float64", TOL64), ("mixed", TOL_MIXED)])
def test_accelerator_score_pairs(monkeypatch, precision, tol):
    monkeypatch.setenv("VARDEPTH_CUDA_PRECISION", precision)
    jobs = _str_jobs("grid")
    close(accelerator.score_pairs(jobs), G["grid_dense_m0"], tol)
    close(accelerator.score_pairs(jobs, anchored=True), G["grid_dense_m1"], tol)
    last = accelerator._engine().last
    assert last["backend"] == "hip" and last["pairs"] == len(jobs) and last["precision"] == precision
    assert last["dp_cells"] == sum(len(s) * len(h) for s, q, h in jobs)


def test_accelerator_score_many_region(monkeypatch):
    monkeypatch.setenv("VARDEPTH_CUDA_PRECISION", "mixed")
    jobs = _str_jobs("region")
    R, H = (int(v) for v in G["region_shape"])
    reads = [jobs[i * H][0] for i in range(R)]
    quals = [jobs[i * H][1] for i in range(R)]
    haps = [jobs[j][2] for j in range(H)]
    out = accelerator.score_many(reads, quals, haps)
    assert out.shape == (R, H)
    close(out.ravel(), G["region_dense_m0"], TOL_MIXED)


def test_accelerator_memory_budget_chunks(monkeypatch):
    monkeypatch.setenv("VARDEPTH_CUDA_PRECISION", "float64")
    eng = accelerator.HipPairHMM(memory_mb=8, precision="float64")
    eng.memory_bytes = 20000  # force several launches
    jobs = _str_jobs("grid")
    close(eng.score(jobs), G["grid_dense_m0"], TOL64)
    assert eng.last["batches"] > 1


def test_batch_pairs_and_trie_routing():
    batch._CACHE.clear()
    jobs = _str_jobs("reuse")
    for mode in (0, 1, 2):
        # canonical: dense FP64 pairhmm_pairs on the device
        close(batch.pairs(jobs, mode, canonical=True), G["reuse_dense_m%d" % mode], TOL64)
        # trie routing: the underflowing 'G'*600 Q0 job is recomputed by the dense path
        out = batch.pairs(jobs, mode)
        close(out, G["reuse_dense_m%d" % mode], TOL_MIXED)
        assert np.isfinite(out[-1])


def test_trie_vector_score_region():
    jobs = _str_jobs("region")
    out = trie_vector.score(jobs, 0, batch._uncached_pairs)
    close(out, G["region_dense_m0"], TOL_MIXED)


def test_pair_scores_many_region(monkeypatch):
    monkeypatch.setenv("VARDEPTH_PAIRHMM_BACKEND", "cpu")
    batch._CACHE.clear()
    jobs = _str_jobs("region")
    R, H = (int(v) for v in G["region_shape"])
    reads = [jobs[i * H][0] for i in range(R)]
    quals = [jobs[i * H][1] for i in range(R)]
    haps = [jobs[j][2] for j in range(H)]
    out = caller.pair_scores_many(caller.core(), reads, quals, haps)
    assert out.shape == (R, H)
    close(out.ravel(), G["region_dense_m0"], TOL_MIXED)


def test_caller_pair_score_core():
    jobs = _str_jobs("grid")
    lib = caller.core()
    got = [caller.pair_score(lib, s, list(q), h) for s, q, h in jobs]
    close(got, G["grid_core"], TOL64)
    s, q = jobs[0][0], list(jobs[0][1])
    haps = [h for _, _, h in jobs[:6]]
    close(caller.pair_scores(lib, s, q, haps), [caller.pair_score(lib, s, q, h) for h in haps], TOL64)


def test_prepared_wide():
    jobs = _str_jobs("grid")
    got = [whole_pairhmm.Prepared([h])(s, q)[0] for s, q, h in jobs]
    close(got, G["grid_wide"], TOL64)


def test_cpu_pairs_and_simd_pairs_fp64():
    """The reference's exact scorers (accelerator.py:42-62): anchored ->
    anchored_pairhmm's model, else wide read_haplotype's; the SIMD form groups
    haplotypes by read.  Both FP64 on the device against the reference's own
    scalar outputs; empty reads or haplotypes give -inf (test_accelerator.py:33-40)."""
    jobs = _str_jobs("grid")
    close(accelerator.cpu_pairs(jobs), G["grid_wide"], TOL64)
    close(accelerator.cpu_pairs(jobs, anchored=True), G["grid_anch"], TOL64)
    close(accelerator.cpu_simd_pairs(jobs), G["grid_wide"], TOL64)
    # shuffled pairs with shared reads and repeated qualities keep their order
    rng = np.random.default_rng(4)
    perm = rng.permutation(len(jobs))
    shuffled = [jobs[i] for i in perm] + [("", b"", "A"), ("A", b"\x1e", "")]
    out = accelerator.cpu_simd_pairs(shuffled)
    close(out[:-2], G["grid_wide"][perm], TOL64)
    assert np.isneginf(out[-2:]).all()
    assert np.isneginf(accelerator.cpu_pairs([("", [], "A"), ("A", [30], "")])).all()
    assert accelerator.cpu_simd_pairs([]).shape == (0,)
    # the reference test's cases: same read at two qualities, one read against several haplotypes
    pairs = [("ACGT" * 8, [30] * 32, "ACGT" * 9), ("ACGT" * 8, [3] * 32, "ACGT" * 9),
             ("GATTACA", [20] * 7, "GGGATTACAT"), ("ACGT" * 8, [30] * 32, "TCGT" * 9), ("", [], "A"), ("A", [30], "")]
    np.testing.assert_allclose(accelerator.cpu_simd_pairs(pairs), accelerator.cpu_pairs(pairs), atol=1e-11, rtol=0)
    expect = [oracle.dense(s.encode(), bytes(q), h.encode(), 2) for s, q, h in pairs]
    close(accelerator.cpu_pairs(pairs), expect, TOL64)


def test_score_many_matches_prepared():
    """score_many equals Prepared(haps)(read, q) per read (test_accelerator.py:21-27)."""
    reads = ['ACGT' * 8, 'A' * 30, 'GATTACA']
    quals = [[30] * len(x) for x in reads]
    haps = ['TT' + r + 'GG' for r in reads] + ['ACGT' * 10]
    expected = np.array([whole_pairhmm.Prepared(haps, mode=0)(r, q) for r, q in zip(reads, quals)])
    for sel in ("cpu", "hip"):
        np.testing.assert_allclose(accelerator.score_many(reads, quals, haps, selection=sel), expected,
                                   atol=1e-9, rtol=0)
    assert accelerator.score_many([], [], haps, selection='cpu').shape == (0, 4)
    assert accelerator.score_many(reads, quals, [], selection='cpu').shape == (3, 0)


def test_canonical_batch_equals_scalar_anchored():
    """cpu_batch.pairs(jobs, mode=1, canonical=True) is bit-identical to the
    scalar anchored_pairhmm, on repeated calls (test_cpu_catalog.py:22-31);
    here both run FP64 on the device, the batch as one launch of all jobs."""
    rng = np.random.default_rng(7713)
    jobs, expected = [], []
    lib = native.kernel()
    for _ in range(31):
        n = int(rng.integers(12, 140))
        s = ''.join(rng.choice(list('ACGT'), n))
        h = s[:n // 2] + 'TA' + s[n // 2:]
        q = rng.integers(0, 61, n).tolist()
        jobs.append((s, q, h))
        expected.append(lib.anchored_pairhmm(s.encode(), (ctypes.c_ubyte * n)(*q), n, h.encode(), len(h), .001, .1))
    for repeat in range(2):
        batch._CACHE.clear()
        np.testing.assert_array_equal(batch.pairs(jobs, mode=1, canonical=True), expected)
        np.testing.assert_array_equal(batch.pairs(jobs, mode=1, canonical=True), expected)   # from the cache
    # and the same values in another order and inside a larger batch
    batch._CACHE.clear()
    perm = rng.permutation(len(jobs))
    extra = [(''.join(rng.choice(list('ACGT'), 240)), [30] * 240, ''.join(rng.choice(list('ACGT'), 300)))
             for _ in range(9)]
    got = batch.pairs([jobs[i] for i in perm] + extra, mode=1, canonical=True)
    np.testing.assert_array_equal(got[:len(jobs)], np.asarray(expected)[perm])
    ref = [oracle.dense(s.encode(), bytes(q), h.encode(), 1) for s, q, h in jobs]
    close(expected, ref, TOL64)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_scalar_fast_path_equals_batched_fp64(mode):
    """The one-read entry points (one packed copy, one FP64 unit kernel, no
    planner) return the batched FP64 forward's values bit for bit: reads of 1
    to 300 bases (up to 5 tiles of 64 rows), up to 11 haplotypes per call (3
    blocks of 4), empty haplotypes and negative lengths (-inf)."""
    from vardepth_amd import engine, workloads
    rng = np.random.default_rng(60 + mode)
    lib = native.kernel() if mode == 1 else (native.core() if mode == 0 else native.wide_library())
    name = "anchored_pairhmm" if mode == 1 else "read_haplotype"
    for n in (1, 37, 150, 300):
        read = rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes()
        qual = rng.integers(2, 41, n).astype(np.uint8).tobytes()
        haps = [rng.choice(list(b"ACGTN"), int(m)).astype(np.uint8).tobytes() for m in rng.integers(1, 360, 11)]
        haps[3] = b""
        b = workloads.from_regions([([read], [qual], haps)], 0.001, 0.1)
        want = engine.forward(b, mode=mode, precision="float64")
        q = (ctypes.c_ubyte * n).from_buffer_copy(qual)
        one = np.array([getattr(lib, name)(read, q, n, h, len(h), 0.001, 0.1) for h in haps])
        np.testing.assert_array_equal(one, want)
        if mode != 1:
            hp = (ctypes.c_char_p * len(haps))(*haps)
            lens = (ctypes.c_int * len(haps))(*[len(h) for h in haps])
            out = np.empty(len(haps))
            lib.read_haplotype_batch(read, q, n, hp, lens, len(haps), 0.001, 0.1,
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            np.testing.assert_array_equal(out, want)
            lens[5] = -3  # a negative length is an empty haplotype
            lib.read_haplotype_batch(read, q, n, hp, lens, len(haps), 0.001, 0.1,
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            assert out[5] == -np.inf and np.array_equal(np.delete(out, 5), np.delete(want, 5))
