"""Pin the CPU oracle to the reference: golden vectors made by the reference
C++ compiled from source (tests/golden/make_golden.py) must be reproduced
bit for bit (same IEEE operation order, contraction off, same libm)."""
import numpy as np
import pytest

import oracle
from tests.golden import fixtures as fx

G = fx.load()


def eq(actual, expected):
    np.testing.assert_array_equal(np.asarray(actual), np.asarray(expected))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_dense_grid_matches_pairhmm_pairs(mode):
    jobs = fx.pairs("grid")
    eq(oracle.pairs(jobs, mode), G["grid_dense_m%d" % mode])
    # lane width does not change any lane's arithmetic (batch_hmm.cpp:17 vs :134)
    eq(G["grid_dense2_m%d" % mode], G["grid_dense_m%d" % mode])


@pytest.mark.parametrize("fn,mode", [("wide", 2), ("core", 0), ("anch", 1)])
def test_scalar_variants_grid(fn, mode):
    jobs = fx.pairs("grid")
    eq(oracle.pairs(jobs, mode, strict=True), G["grid_" + fn])


@pytest.mark.parametrize("shape", range(6))
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_trie_grid(shape, mode):
    jobs = fx.pairs("grid")[6 * shape:6 * shape + 6]
    out, cells = oracle.trie_many([j[0] for j in jobs], [j[1] for j in jobs], [j[2] for j in jobs], mode)
    eq(out, G["grid_trie%d_m%d" % (shape, mode)])
    assert cells == int(G["grid_trie%d_m%d_cells" % (shape, mode)][0])


def test_parametric_gap_ext():
    jobs = fx.pairs("param")
    for gi, g in enumerate(G["param_gaps"]):
        for xi, x in enumerate(G["param_exts"]):
            eq(oracle.pairs(jobs, 2, g, x, strict=True), G["param_wide_g%d_x%d" % (gi, xi)])
            eq(oracle.pairs(jobs, 0, g, x, strict=True), G["param_core_g%d_x%d" % (gi, xi)])
            eq(oracle.pairs(jobs, 1, g, x, strict=True), G["param_anchored_g%d_x%d" % (gi, xi)])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_edge_cases(mode):
    jobs = fx.pairs("edge")
    eq(oracle.pairs(jobs, mode), G["edge_dense_m%d" % mode])
    reads = [j for j in jobs if j[0]]
    out, cells = oracle.trie_many([j[0] for j in reads], [j[1] for j in reads], fx.strings("edge_trie_haps"), mode)
    eq(out, G["edge_trie_m%d" % mode])
    assert cells == int(G["edge_trie_m%d_cells" % mode][0])


def test_edge_scalar_variants():
    jobs = fx.pairs("edge")
    eq(oracle.pairs(jobs, 2, strict=True), G["edge_wide"])
    eq(oracle.pairs(jobs, 0, strict=True), G["edge_core"])
    eq(oracle.pairs(jobs, 1, strict=True), G["edge_anch"])
    # empty read or haplotype -> -inf (batch_hmm.cpp:70)
    assert np.isneginf(G["edge_dense_m0"][:2]).all()


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_active_region(mode):
    jobs = fx.pairs("region")
    R, H = (int(v) for v in G["region_shape"])
    reads = [jobs[i * H][0] for i in range(R)]
    quals = [jobs[i * H][1] for i in range(R)]
    haps = [jobs[j][2] for j in range(H)]
    out, cells = oracle.trie_many(reads, quals, haps, mode)
    eq(out, G["region_trie_m%d" % mode])
    assert cells == int(G["region_trie_m%d_cells" % mode][0])
    eq(oracle.pairs(jobs, mode), G["region_dense_m%d" % mode])


def test_active_region_q45():
    jobs = fx.pairs("region")
    gap, ext = fx.meta()["sets"]["region"]["gap"], fx.meta()["sets"]["region"]["ext"]
    eq(oracle.pairs(jobs, 2, gap, ext, strict=True), G["region_wide_q45"])
    eq(oracle.pairs(jobs, 1, gap, ext, strict=True), G["region_anch_q45"])
    eq(oracle.pairs(jobs, 0, gap, ext, strict=True), G["region_core_q45"])


def test_precision_stress():
    jobs = fx.pairs("stress")
    for mode in range(3):
        eq(oracle.pairs(jobs, mode), G["stress_dense_m%d" % mode])
    eq(oracle.pairs(jobs, 2, 10 ** -4.5, 0.1, strict=True), G["stress_wide_q45"])
    eq(oracle.pairs(jobs, 1, 10 ** -4.5, 0.1, strict=True), G["stress_anch_q45"])


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("lanes", [0, 4, 2])
def test_trie_shared_suffix(mode, lanes):
    jobs = fx.pairs("suffix")
    haps = fx.strings("suffix_haps")
    out, cells = oracle.trie_many([j[0] for j in jobs], [j[1] for j in jobs], haps, mode, lanes)
    eq(out, G["suffix_trie_m%d_l%d" % (mode, lanes)])
    assert cells == int(G["suffix_trie_m%d_l%d_cells" % (mode, lanes)][0])


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_prefix_reuse_jobs(mode):
    jobs = fx.pairs("reuse")
    eq(oracle.pairs(jobs, mode), G["reuse_dense_m%d" % mode])
    vals = [oracle.trie_many([s], [q], [h], mode)[0][0, 0] for s, q, h in jobs]
    eq(vals, G["reuse_trie_m%d" % mode])
    # the 'G'*600 Q0 case underflows the trie (NaN) -> dense fallback path
    assert np.isnan(G["reuse_trie_m%d" % mode][-1])
    # trie agrees with dense where it is finite (test_cpu_reuse.py:21 tolerance)
    t, d = G["reuse_trie_m%d" % mode], G["reuse_dense_m%d" % mode]
    ok = np.isfinite(t)
    np.testing.assert_allclose(t[ok], d[ok], atol=1e-9, rtol=1e-12)
