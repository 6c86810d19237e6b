"""Host logic of the reference-interface mirrors, on the CPU.

The device engine is replaced by a stub that evaluates each pair with the
oracle, so these tests check validation, caching, grouping, NaN routing,
order restoration and dispatch -- the behaviour of accelerator.py:25-40,
cpu_batch.py:20-66, trie_vector.py:16-42 and base/caller.py:151-160 -- without
a GPU.  The stub is test scaffolding; the product path has no CPU engine."""
import math

import numpy as np
import pytest

import oracle
from vardepth_amd import accelerator, batch, caller, engine, trie_vector, whole_pairhmm, workloads

RNG = np.random.default_rng(11)


def rand_dna(n):
    return "".join(RNG.choice(list("ACGT"), n))


def rand_q(n, lo=10, hi=40):
    return bytes(RNG.integers(lo, hi + 1, n).astype(np.uint8))


class OracleEngine:
    """engine.forward stand-in: the oracle's dense value of every pair, in the
    batch's output layout; records the batches it saw."""

    def __init__(self):
        self.calls = []

    def __call__(self, b, mode=0, precision="mixed", err=None, stats=False):
        self.calls.append((b, mode, precision))
        out = []
        for k in range(b.nregions):
            reads, quals, haps = b.region(k)
            out += [oracle.dense(s, q, h, mode, b.gap, b.ext) for s, q in zip(reads, quals) for h in haps]
        out = np.array(out, np.float64)
        st = dict(device_ms=0.0, fwd64_ms=0.0, fp64_pairs=0, flagged=0)
        return (out, st) if stats else out


@pytest.fixture
def stub_engine(monkeypatch):
    eng = OracleEngine()
    monkeypatch.setattr(engine, "forward", eng)
    return eng


# --- accelerator.validate_pairs (accelerator.py:25-40) -------------------------------------

def test_validate_pairs_accepts_and_converts():
    q_list = [30, 31, 32]
    q_arr = np.array([1, 2, 3], dtype=np.int64)
    out = accelerator.validate_pairs([("ACG", b"\x05\x06\x07", "ACGT"), ("ACN", q_list, "A"), ("AAA", q_arr, "")])
    assert [o[0] for o in out] == ["ACG", "ACN", "AAA"]
    for o in out:
        assert o[1].dtype == np.uint8
    assert out[1][1].tolist() == q_list
    # float qualities with integer values are accepted (np.floor check)
    accelerator.validate_pairs([("AC", np.array([3.0, 4.0]), "A")])


@pytest.mark.parametrize("pair,msg", [
    (("ACG", b"\x05\x06", "A"), "Read/quality lengths differ"),
    (("acg", b"\x05\x06\x07", "A"), "Uppercase DNA required"),
    (("ACG", b"\x05\x06\x07", "ACGX"), "Uppercase DNA required"),
    ((b"ACG", b"\x05\x06\x07", "A"), "Uppercase DNA required"),
    (("ACG", [1, 2, 256], "A"), "integer Phred"),
    (("ACG", [1, -1, 2], "A"), "integer Phred"),
    (("ACG", [1.5, 2, 3], "A"), "integer Phred"),
    (("ACG", [float("nan"), 2, 3], "A"), "integer Phred"),
    (("AC", np.zeros((1, 2)), "A"), "Read/quality lengths differ"),
])
def test_validate_pairs_rejects(pair, msg):
    with pytest.raises(ValueError, match=msg):
        accelerator.validate_pairs([pair])


def test_score_pairs_empty_and_reshape_without_device():
    assert accelerator.score_pairs([]).shape == (0,)
    assert accelerator.score_many([], [], ["ACGT", "AC"]).shape == (0, 2)
    assert accelerator.score_many(["ACGT"], [b"\x1e" * 4], []).shape == (1, 0)
    with pytest.raises(ValueError, match="batch lengths differ"):
        accelerator.score_many(["A"], [], ["A"])
    with pytest.raises(ValueError, match="Invalid PairHMM backend"):
        accelerator.score_pairs([("A", b"\x1e", "A")], selection="tpu")


def test_backend_env(monkeypatch):
    monkeypatch.setenv("VARDEPTH_PAIRHMM_BACKEND", "bogus")
    with pytest.raises(ValueError):
        accelerator.backend()
    for v in ("cpu", "cuda", "auto", "hip"):
        monkeypatch.setenv("VARDEPTH_PAIRHMM_BACKEND", v)
        assert accelerator.backend() == v


def test_engine_params_validated():
    with pytest.raises(ValueError, match="float64 or mixed"):
        accelerator.HipPairHMM(precision="half")
    with pytest.raises(ValueError, match="at least 8 MiB"):
        accelerator.HipPairHMM(memory_mb=4, precision="mixed")


def test_regions_by_read_groups_pairs():
    q1, q2 = np.array([30, 30], np.uint8), np.array([20, 20], np.uint8)
    pairs = [("AC", q1, "A"), ("AC", q2, "C"), ("AC", q1, "G"), ("GG", q1, "T")]
    regions, slots = accelerator._regions_by_read(pairs, range(4))
    assert slots == [[0, 2], [1], [3]]
    assert regions[0][2] == ["A", "G"]


# --- batch.pairs (cpu_batch.py:20-66) ---------------------------------------------------------

def test_batch_pairs_cache_and_dedup(monkeypatch, stub_engine):
    batch._CACHE.clear()
    s, q = rand_dna(40), rand_q(40)
    haps = [rand_dna(50) for _ in range(3)]
    jobs = [(s, q, h) for h in haps] + [(s, q, haps[0])]
    seen = []
    real = trie_vector.score

    def spy(missing, mode, dense, lanes=0):
        seen.append(list(missing))
        return real(missing, mode, dense, lanes=lanes)

    monkeypatch.setattr(trie_vector, "score", spy)
    out = batch.pairs(jobs, mode=0)
    assert len(seen) == 1 and len(seen[0]) == 3  # the duplicate job is computed once
    expect = [oracle.dense(s.encode(), q, h.encode(), 0) for h in haps]
    np.testing.assert_allclose(out, expect + [expect[0]], rtol=0, atol=1e-12)
    out2 = batch.pairs(jobs[::-1], mode=0)  # all cached, order follows the input
    assert len(seen) == 1
    np.testing.assert_array_equal(out2, out[::-1])
    batch.pairs(jobs, mode=1)  # the mode is part of the key
    assert len(seen) == 2
    batch.pairs(jobs, mode=0, lanes=4)  # so is the lane request
    assert len(seen) == 3


def test_batch_canonical_goes_dense(monkeypatch):
    batch._CACHE.clear()
    calls = []

    def dense(jobs, mode, lanes):
        calls.append((len(jobs), mode, lanes))
        return np.arange(len(jobs), dtype=np.float64) - 7.0

    def no_trie(*a, **k):
        raise AssertionError("canonical=True must not use the trie routing")

    monkeypatch.setattr(batch, "_uncached_pairs", dense)
    monkeypatch.setattr(trie_vector, "score", no_trie)
    jobs = [(rand_dna(5), rand_q(5), rand_dna(6)) for _ in range(3)]
    out = batch.pairs(jobs, mode=2, lanes=8, canonical=True)
    assert calls == [(3, 2, 8)]
    np.testing.assert_array_equal(out, [-7.0, -6.0, -5.0])


def test_batch_lru_eviction(monkeypatch):
    batch._CACHE.clear()
    monkeypatch.setattr(batch, "_CACHE_LIMIT", 4)
    monkeypatch.setattr(batch, "_uncached_pairs", lambda jobs, mode, lanes: np.zeros(len(jobs)))
    jobs = [("A" * (i + 1), b"\x1e" * (i + 1), "A", ) for i in range(6)]
    batch.pairs(jobs, canonical=True)
    assert len(batch._CACHE) == 4
    keys = [k[2] for k in batch._CACHE]
    assert keys == ["A" * 3, "A" * 4, "A" * 5, "A" * 6]  # oldest evicted first


def test_batch_length_mismatch():
    with pytest.raises(ValueError, match="Read and quality lengths differ"):
        batch.pairs([("ACG", b"\x01", "A")])
    with pytest.raises(ValueError, match="Read batch lengths differ"):
        batch.many(["A"], [], ["A"])
    assert batch.pairs([]).shape == (0,)


# --- trie_vector.score (trie_vector.py:16-42) ---------------------------------------------

def test_trie_score_groups_by_hap_set(stub_engine):
    haps_a = [rand_dna(60), rand_dna(45)]
    haps_b = [rand_dna(30)]
    reads = [(rand_dna(n), rand_q(n)) for n in (50, 20, 35)]
    jobs = []
    for s, q in reads[:2]:
        jobs += [(s, q, h) for h in haps_a]
    jobs += [(reads[2][0], reads[2][1], haps_b[0])]
    jobs = jobs[::-1]
    out = trie_vector.score(jobs, 0, lambda *a: pytest.fail("no fallback expected"))
    (b, mode, _), = stub_engine.calls
    assert b.nregions == 2  # two distinct sorted haplotype sets -> two regions, one launch
    r0 = b.region(0)
    assert [len(s) for s in r0[0]] == sorted(len(s) for s in r0[0])  # reads sorted by length
    assert r0[2] == sorted(r0[2])  # haplotypes sorted
    expect = [oracle.dense(s.encode(), q, h.encode(), 0) for s, q, h in jobs]
    np.testing.assert_allclose(out, expect, rtol=0, atol=1e-12)


def test_trie_score_nan_routes_to_dense(stub_engine):
    # 'G'*600 at Q0 against 'G'*601: e = 1 makes every match emit 0, L <= 1e-200
    tiny = ("G" * 600, bytes(600), "G" * 601)
    ok = (rand_dna(30), rand_q(30), rand_dna(40))
    empty = ("", b"", "ACGT")
    routed = []

    def dense(jobs, mode, lanes):
        routed.append(jobs)
        return np.full(len(jobs), -1234.5)

    out = trie_vector.score([ok, tiny, empty], 0, dense)
    assert routed == [[tiny]]
    assert out[1] == -1234.5
    assert np.isneginf(out[2])  # empty inputs stay -inf, not NaN
    assert math.isclose(out[0], oracle.dense(ok[0].encode(), ok[1], ok[2].encode(), 0), abs_tol=1e-12)


# --- caller.pair_scores_many dispatch (base/caller.py:151-160) --------------------------------

def test_pair_scores_many_dispatch(monkeypatch):
    calls = []
    monkeypatch.setattr(accelerator, "score_many", lambda s, q, h: calls.append("accel") or np.zeros((len(s), len(h))))
    monkeypatch.setattr(batch, "many", lambda s, q, h, mode=0: calls.append(("batch", mode)) or np.zeros((len(s), len(h))))
    small = (["ACGT"], [b"\x1e" * 4], ["ACGT", "ACGA"])
    monkeypatch.setenv("VARDEPTH_PAIRHMM_BACKEND", "cpu")
    assert caller.pair_scores_many(None, *small).shape == (1, 2)
    monkeypatch.setenv("VARDEPTH_PAIRHMM_BACKEND", "auto")
    caller.pair_scores_many(None, *small)
    big = (["A" * 2000], [b"\x1e" * 2000], ["A" * 1000])
    caller.pair_scores_many(None, *big)
    monkeypatch.setenv("VARDEPTH_PAIRHMM_BACKEND", "cuda")
    caller.pair_scores_many(None, *small)
    assert calls == [("batch", 0), ("batch", 0), "accel", "accel"]


def test_prepared_lane_request():
    with pytest.raises(ValueError, match="Unsupported lane request"):
        whole_pairhmm.Prepared(["ACGT"], mode=3)


def test_from_regions_layout():
    b = workloads.from_regions([(["AC", "GGT"], [b"\x01\x02", b"\x03\x04\x05"], ["A", "CC", "G"]),
                                (["T"], [b"\x09"], ["TT"])])
    assert b.npairs == 7 and b.cells() == 2 * 4 + 3 * 4 + 2
    assert b.out_off().tolist() == [0, 6, 7]


def test_grad_weights_checked_before_device():
    """A weight vector of the wrong length is refused on the host (reads of
    any length are accepted: those beyond 256 rows take the tiled FP64
    adjoint)."""
    b1 = workloads.from_regions([(["A" * 300], [b"\x1e" * 300], ["A" * 310])])
    with pytest.raises(ValueError, match="one entry per"):
        engine.grad(b1, weights=[1.0, 2.0])


def test_tuning_get_set_validation():
    from vardepth_amd import _lib, autotune
    t = autotune.get()
    assert list(t.fwd_unroll)[1:] == [2] * 16 or all(v in (2, 4) for v in list(t.fwd_unroll)[1:])
    bad = _lib.VdTuning.from_buffer_copy(t)
    bad.fwd_unroll[5] = 3
    with pytest.raises(_lib.VdhmmError, match="fwd_unroll"):
        autotune.set_tuning(bad)
    bad = _lib.VdTuning.from_buffer_copy(t)
    bad.grad_variant[7] = 2
    with pytest.raises(_lib.VdhmmError, match="grad_variant"):
        autotune.set_tuning(bad)
    bad = _lib.VdTuning.from_buffer_copy(t)
    bad.fwd_share[3] = 4
    with pytest.raises(_lib.VdhmmError, match="fwd_share"):
        autotune.set_tuning(bad)
    for chunk in (0, 65):
        bad = _lib.VdTuning.from_buffer_copy(t)
        bad.share_chunk = chunk
        with pytest.raises(_lib.VdhmmError, match="share_chunk"):
            autotune.set_tuning(bad)
    assert t.share_chunk == 4 and list(t.fwd_share)[1:] == [2] * 16   # defaults
    good = _lib.VdTuning.from_buffer_copy(t)
    good.fwd_unroll[10] = 4
    good.fwd_share[10] = 0
    good.share_chunk = 16
    autotune.set_tuning(good)
    assert autotune.get().fwd_unroll[10] == 4 and autotune.get().fwd_share[10] == 0
    assert autotune.get().share_chunk == 16
    autotune.set_tuning(t)


def test_haplotype_length_limit_checked_before_the_device():
    """The longest haplotype the kernels hold in LDS (160 KiB per workgroup on
    gfx950) is known without a device, and a batch above it is refused before
    any device call."""
    from vardepth_amd import _lib
    lim = _lib.lib().vd_max_hap_len(0)
    glim = _lib.lib().vd_max_hap_len(1)
    assert 19000 < glim <= lim < 21000
    b = workloads.from_regions([(["ACGT"], [b"\x1e" * 4], ["A" * (lim + 1)])])
    with pytest.raises(_lib.VdhmmError, match="exceeds the %d bases" % lim):
        engine.forward(b)
    g = workloads.from_regions([(["ACGT"], [b"\x1e" * 4], ["A" * (glim + 1)])])
    with pytest.raises(_lib.VdhmmError, match="gradient kernels"):
        engine.grad(g)


def test_tuning_save_load_resets_absent_fields(tmp_path):
    """autotune.load applies every saved field and gives the ones an entry
    lacks their defaults, the stream layout included (vd_tuning.class_overlap)."""
    import json
    from vardepth_amd import _lib, autotune
    base = autotune.get()
    try:
        t = _lib.VdTuning.from_buffer_copy(base)
        t.class_overlap = 0
        t.share_chunk = 16
        t.fwd_share[7] = 1
        autotune.set_tuning(t)
        assert _lib.lib().vd_class_overlap(-1) == 0 and autotune.get().class_overlap == 0
        path = tmp_path / "tune.json"
        autotune.save(path, "a")
        saved = json.loads(path.read_text())["a"]
        assert saved["class_overlap"] == 0 and saved["share_chunk"] == 16
        # an entry without the stream layout and the chunk: both return to their defaults
        del saved["class_overlap"], saved["share_chunk"]
        path.write_text(json.dumps({"b": saved}))
        assert autotune.load(path, "b")
        got = autotune.get()
        assert got.class_overlap == 1 and got.share_chunk == 4 and got.fwd_share[7] == 1
        assert not autotune.load(path, "missing")
        bad = _lib.VdTuning.from_buffer_copy(base)
        bad.class_overlap = 2
        with pytest.raises(_lib.VdhmmError, match="class_overlap"):
            autotune.set_tuning(bad)
    finally:
        autotune.set_tuning(base)
