"""Write tests/golden/golden_pairhmm.npz from the reference compiled from source.

Run in a container that holds the reference tree:

    make -C oracle ref          # compiles the reference C++ into oracle/_ref
    python tests/golden/make_golden.py

The reference sources are compiled by oracle/Makefile with the reference's
own flags; nothing of the reference is stored here -- only inputs and the
outputs its functions produced on them:

* pairhmm_pairs  (batch_hmm.cpp:192)      modes 0/1/2, lanes 0 and 2
* trie_many      (trie_vector.cpp:284)    modes 0/1/2, with the cells counter
* read_haplotype (wide.cpp:9)             mode-2 scalar, parametric gap/ext
* read_haplotype (base/core/core.cpp:9)   mode-0 scalar, parametric gap/ext
* anchored_pairhmm (anchored.cpp:5)       mode-1 scalar, parametric gap/ext

Case sets: the reference's CUDA-test shape grid (test_accelerator.py:44),
a gap/ext grid, edge cases (empty, Q0, q>60, N and other bytes, the
'G'*600/Q0 NaN case of test_cpu_reuse.py:19), one 64x8 active region, 8
precision-stress pairs, a trie shared-suffix set, and the prefix-reuse jobs
of test_cpu_reuse.py:15-19.
"""
from __future__ import annotations

import json
import platform
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from vardepth_amd import workloads  # noqa: E402

OUT = Path(__file__).resolve().parent / "golden_pairhmm.npz"


def pack(prefix, pairs, store):
    reads = [p[0].encode("latin-1") if isinstance(p[0], str) else bytes(p[0]) for p in pairs]
    quals = [oracle._qual(p[1]) for p in pairs]
    haps = [p[2].encode("latin-1") if isinstance(p[2], str) else bytes(p[2]) for p in pairs]
    store[prefix + "_reads"] = np.frombuffer(b"".join(reads), np.uint8)
    store[prefix + "_quals"] = np.frombuffer(b"".join(quals), np.uint8)
    store[prefix + "_read_off"] = np.concatenate([[0], np.cumsum([len(r) for r in reads])]).astype(np.int64)
    store[prefix + "_haps"] = np.frombuffer(b"".join(haps), np.uint8)
    store[prefix + "_hap_off"] = np.concatenate([[0], np.cumsum([len(h) for h in haps])]).astype(np.int64)


def rand_seq(rng, n):
    return "".join(rng.choice(list("ACGT"), n))


def main():
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref` with the reference present")
    store = {}
    meta = {"sets": {}}

    # 1. shape grid of the reference CUDA parity test (test_accelerator.py:43-47)
    rng = np.random.default_rng(31)
    grid = []
    shapes = [(1, 1), (7, 9), (30, 35), (151, 201), (300, 450), (700, 750)]
    for n, m in shapes:
        for _ in range(6):
            grid.append((rand_seq(rng, n), rng.integers(0, 61, n).astype(np.uint8), rand_seq(rng, m)))
    pack("grid", grid, store)
    for mode in range(3):
        store["grid_dense_m%d" % mode] = ref.pairhmm_pairs(grid, mode, 0)
        store["grid_dense2_m%d" % mode] = ref.pairhmm_pairs(grid, mode, 2)
    store["grid_wide"] = ref.scalar("wide", grid)
    store["grid_core"] = ref.scalar("core", grid)
    store["grid_anch"] = ref.scalar("anchored", grid)
    # trie per shape: 6 reads x 6 haplotypes
    for s, (n, m) in enumerate(shapes):
        sub = grid[6 * s:6 * s + 6]
        for mode in range(3):
            out, cells = ref.trie_many([p[0] for p in sub], [p[1] for p in sub], [p[2] for p in sub], mode)
            store["grid_trie%d_m%d" % (s, mode)] = out
            store["grid_trie%d_m%d_cells" % (s, mode)] = np.array([cells], np.uint64)
    meta["sets"]["grid"] = {"shapes": shapes, "per_shape": 6, "seed": 31}

    # 2. gap/ext grid through the parametric scalar entry points
    rng = np.random.default_rng(4501)
    par = []
    for n, m in [(7, 9), (30, 35), (60, 61), (151, 201)]:
        s = rand_seq(rng, n)
        h = s[: n // 2] + rand_seq(rng, 2) + s[n // 2:] + rand_seq(rng, max(0, m - n - 2))
        par.append((s, rng.integers(2, 41, n).astype(np.uint8), h[:m]))
    pack("param", par, store)
    gaps, exts = [1e-3, 10 ** -4.5, 1e-2], [0.05, 0.1, 0.3]
    for gi, g in enumerate(gaps):
        for xi, x in enumerate(exts):
            for fn in ("wide", "core", "anchored"):
                store["param_%s_g%d_x%d" % (fn, gi, xi)] = ref.scalar(fn, par, g, x)
    store["param_gaps"] = np.array(gaps)
    store["param_exts"] = np.array(exts)

    # 3. edge cases
    edge = [("", [], "A"), ("A", [30], ""), ("A", [30], "A"), ("A", [30], "C"), ("A", [0], "A"),
            ("G" * 600, [0] * 600, "G" * 601), ("ACGT" * 10, [60] * 40, "ACGT" * 12),
            ("ACGT" * 10, [61] * 40, "ACGT" * 12), ("ACGT" * 10, [255] * 40, "ACGT" * 12),
            ("ACNTNNAC", [20] * 8, "GACNTNNACG"), ("ACNTGGAC", [20] * 8, "GACATGGACG"),
            ("ACXTaGAC", [25] * 8, "GACXTaGACG"), ("ACGTAGAC", [25] * 8, "GACXTaGACG"),
            ("A" * 100, [60] * 100, "A" * 100), ("A" * 200, [35] * 200, "A" * 450),
            ("ACGT" * 100, [60] * 400, "TGCA" * 160), ("C" * 30, [2] * 30, "G" * 40),
            ("T", [40], "ACGTACGTACGT"), ("ACGTACGTACGT", [40] * 12, "T")]
    pack("edge", edge, store)
    for mode in range(3):
        store["edge_dense_m%d" % mode] = ref.pairhmm_pairs(edge, mode, 0)
    store["edge_wide"] = ref.scalar("wide", edge)
    store["edge_core"] = ref.scalar("core", edge)
    store["edge_anch"] = ref.scalar("anchored", edge)
    # trie over the non-empty edge reads against a few haplotypes each
    ereads = [p for p in edge if p[0]]
    ehaps = ["GACNTNNACG", "GACXTaGACG", "ACGT" * 12, "A" * 100, "G" * 601, "T"]
    store["edge_trie_haps"] = np.frombuffer("|".join(ehaps).encode(), np.uint8)
    for mode in range(3):
        out, cells = ref.trie_many([p[0] for p in ereads], [p[1] for p in ereads], ehaps, mode)
        store["edge_trie_m%d" % mode] = out
        store["edge_trie_m%d_cells" % mode] = np.array([cells], np.uint64)

    # 4. one active region (64 reads x 8 haplotypes of 200-400 bp)
    b = workloads.single_region(seed=41001)
    reads, quals, haps = b.region(0)
    region = [(s, q, h) for s, q in zip(reads, quals) for h in haps]
    pack("region", region, store)
    store["region_shape"] = np.array([len(reads), len(haps)], np.int64)
    for mode in range(3):
        out, cells = ref.trie_many(reads, quals, haps, mode)
        store["region_trie_m%d" % mode] = out
        store["region_trie_m%d_cells" % mode] = np.array([cells], np.uint64)
        store["region_dense_m%d" % mode] = ref.pairhmm_pairs(region, mode, 0)
    store["region_wide_q45"] = ref.scalar("wide", region, b.gap, b.ext)
    store["region_anch_q45"] = ref.scalar("anchored", region, b.gap, b.ext)
    store["region_core_q45"] = ref.scalar("core", region, b.gap, b.ext)
    meta["sets"]["region"] = {"seed": 41001, "gap": b.gap, "ext": b.ext}

    # 5. precision stress: 8 pairs of 250 bp Phred 2-15 reads vs 1-2 kb haplotypes
    sb = workloads.precision_stress(seed=41003, regions=2, reads=4)
    stress = []
    for k in range(sb.nregions):
        rr, qq, hh = sb.region(k)
        for i in range(4):
            stress.append((rr[i], qq[i], hh[i % len(hh)]))
    stress = stress[:8]
    pack("stress", stress, store)
    for mode in range(3):
        store["stress_dense_m%d" % mode] = ref.pairhmm_pairs(stress, mode, 0)
    store["stress_wide_q45"] = ref.scalar("wide", stress, sb.gap, sb.ext)
    store["stress_anch_q45"] = ref.scalar("anchored", stress, sb.gap, sb.ext)

    # 6. trie shared-suffix continuation (trie_vector.cpp:198-257)
    rng = np.random.default_rng(2571)
    tail = rand_seq(rng, 40)
    stem = rand_seq(rng, 100)
    shaps = sorted({stem + tail, stem[:30] + "A" + stem[30:] + tail, stem[:45] + stem[47:] + tail,
                    rand_seq(rng, 95) + tail, stem[:10] + "T" + stem[11:] + tail})
    sreads = []
    for n in [20, 35, 50, 64, 75, 80, 90, 95, 100, 110, 117]:
        src = shaps[int(rng.integers(0, len(shaps)))]
        o = int(rng.integers(0, len(src) - n + 1))
        sreads.append((src[o:o + n], rng.integers(5, 41, n).astype(np.uint8)))
    pack("suffix", [(s, q, "") for s, q in sreads], store)
    store["suffix_haps"] = np.frombuffer("|".join(shaps).encode(), np.uint8)
    for mode in range(3):
        for lanes in (0, 4, 2):
            out, cells = ref.trie_many([s for s, q in sreads], [q for s, q in sreads], shaps, mode, lanes)
            store["suffix_trie_m%d_l%d" % (mode, lanes)] = out
            store["suffix_trie_m%d_l%d_cells" % (mode, lanes)] = np.array([cells], np.uint64)

    # 7. prefix-reuse jobs of test_cpu_reuse.py:15-19 through trie (per read, sorted haplotypes)
    rng = np.random.default_rng(441901)
    jobs = []
    for n in [1, 3, 30, 151]:
        s = rand_seq(rng, n)
        q = rng.integers(0, 61, n).astype(np.uint8)
        for h in [s, s[: n // 2] + "T" + s[n // 2:], s[: max(1, n - 2)], "C" * n]:
            jobs.append((s, q, h))
    jobs.extend([("", np.zeros(0, np.uint8), "A"), ("A", np.array([30], np.uint8), ""),
                 ("G" * 600, np.zeros(600, np.uint8), "G" * 601)])
    pack("reuse", jobs, store)
    for mode in range(3):
        store["reuse_dense_m%d" % mode] = ref.pairhmm_pairs(jobs, mode, 0)
        vals = np.empty(len(jobs))
        for i, (s, q, h) in enumerate(jobs):
            out, _ = ref.trie_many([s], [q], [h], mode)
            vals[i] = out[0, 0]
        store["reuse_trie_m%d" % mode] = vals

    cxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    meta["provenance"] = {
        "compiler": cxx, "glibc": platform.libc_ver()[1], "machine": platform.machine(),
        "flags": "-O3 -std=c++17 [-ffp-contract=off] -shared -fPIC (core.cpp: -std=c++20)",
        "source": "caller/native/caller/{batch_hmm,trie_vector,wide,anchored}.cpp, base/core/core.cpp",
    }
    store["meta_json"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    np.savez_compressed(OUT, **store)
    print("wrote", OUT, OUT.stat().st_size, "bytes,", len(store), "arrays")


if __name__ == "__main__":
    main()
