"""Write tests/golden/golden_mixture.npz from the reference compiled from source.

Run in a container that holds the reference tree:

    make -C oracle ref                    # mixture_two.cpp, mixture_general.cpp
    python tests/golden/make_golden_mixture.py

oracle/Makefile compiles the two reference sources with the reference's own
flags (native_build.py:24-36).  Only inputs and the outputs the reference
functions produced on them are stored:

* mixture_general (mixture_general.cpp:9)  triangular and square, priors
  none / 0 / 1e-3 / 0.3 / 0.999, n = 1..64 haplotypes, 0..150 rows, with
  zero entries (the 1e-300 floor), repeated rows (the row cache), a row of
  equal columns, huge and tiny magnitudes;
* mixture_two (mixture_two.cpp:9)          the same kinds for two columns.
"""
from __future__ import annotations

import platform
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent / "golden_mixture.npz"
PRIORS = [-1.0, 0.0, 0.001, 0.3, 0.999]


def matrices(rng):
    shapes = [(0, 1), (0, 3), (1, 1), (3, 1), (1, 2), (5, 2), (40, 2), (2, 3), (17, 4), (9, 5),
              (60, 6), (33, 8), (90, 11), (25, 16), (12, 23), (150, 7), (20, 32), (6, 64)]
    for rows, n in shapes:
        m = rng.random((rows, n)) ** 3
        if rows >= 3:
            m[1, rng.integers(n)] = 0.0             # floor at 1e-300
            m[2] = m[0]                             # repeated row
        if rows >= 5:
            m[3] = 0.5                              # equal columns
            m[4, 0] = 1e-200                        # tiny, and a larger column
        yield m
    # heavy rows: normalised molecule rows (max 1) with a near-zero tail
    for rows, n in [(50, 3), (80, 9)]:
        m = np.exp(-rng.exponential(8.0, (rows, n)))
        m /= m.max(axis=1, keepdims=True)
        yield m
    # large magnitudes
    yield rng.random((4, 3)) * 1e250


def main():
    ref = oracle.ref()
    if ref is None or ref.mix_general is None or ref.mix_two is None:
        sys.exit("build the reference first: make -C oracle ref")
    rng = np.random.default_rng(41001)
    store = {}
    mats = list(matrices(rng))
    g_in, g_shape, g_prior, g_tri, g_out, g_in_off, g_out_off = [], [], [], [], [], [0], [0]
    for m in mats:
        for prior in PRIORS:
            for tri in (1, 0):
                if not tri and prior not in (-1.0, 0.001):
                    continue
                out = ref.mixture_general(m, None if prior < 0 else prior, tri)
                g_in.append(m.ravel())
                g_shape.append(m.shape)
                g_prior.append(prior)
                g_tri.append(tri)
                g_out.append(out)
                g_in_off.append(g_in_off[-1] + m.size)
                g_out_off.append(g_out_off[-1] + out.size)
    store["general_in"] = np.concatenate(g_in)
    store["general_in_off"] = np.array(g_in_off, np.int64)
    store["general_shape"] = np.array(g_shape, np.int64)
    store["general_prior"] = np.array(g_prior)
    store["general_tri"] = np.array(g_tri, np.int8)
    store["general_out"] = np.concatenate(g_out)
    store["general_out_off"] = np.array(g_out_off, np.int64)

    t_in, t_rows, t_prior, t_out = [], [], [], []
    for m in [x for x in mats if x.shape[1] == 2] + [rng.random((70, 2)) ** 4]:
        for prior in (0.0, 0.001, 0.3, 0.999):
            t_in.append(m.ravel())
            t_rows.append(len(m))
            t_prior.append(prior)
            t_out.append(ref.mixture_two(m, prior).ravel())
    store["two_in"] = np.concatenate(t_in)
    store["two_rows"] = np.array(t_rows, np.int64)
    store["two_prior"] = np.array(t_prior)
    store["two_out"] = np.stack(t_out)
    store["meta"] = np.array(f"reference mixture_two.cpp / mixture_general.cpp built by oracle/Makefile "
                             f"on {platform.machine()}; numpy {np.__version__}")
    np.savez_compressed(OUT, **store)
    print("wrote", OUT, sum(v.nbytes for v in store.values()), "bytes")


if __name__ == "__main__":
    main()
