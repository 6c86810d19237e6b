"""Loader for the golden vectors written by make_golden.py."""
from __future__ import annotations

import functools
import json
from pathlib import Path

import numpy as np

PATH = Path(__file__).resolve().parent / "golden_pairhmm.npz"


@functools.lru_cache(maxsize=1)
def load():
    with np.load(PATH, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def pairs(prefix):
    """List of (read bytes, qual bytes, hap bytes) of a stored case set."""
    g = load()
    rb, qb, ro = g[prefix + "_reads"], g[prefix + "_quals"], g[prefix + "_read_off"]
    hb, ho = g[prefix + "_haps"], g[prefix + "_hap_off"]
    out = []
    for i in range(len(ro) - 1):
        out.append((rb[ro[i]:ro[i + 1]].tobytes(), qb[ro[i]:ro[i + 1]].tobytes(),
                    hb[ho[i]:ho[i + 1]].tobytes()))
    return out


def strings(key):
    return [s.encode() for s in load()[key].tobytes().decode().split("|")]


def meta():
    return json.loads(load()["meta_json"].tobytes().decode())


MIXTURE_PATH = Path(__file__).resolve().parent / "golden_mixture.npz"


@functools.lru_cache(maxsize=1)
def load_mixture():
    with np.load(MIXTURE_PATH, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def mixture_general_cases():
    """(matrix, prior or None, triangular, expected) of mixture_general.cpp:9."""
    g = load_mixture()
    out = []
    for i in range(len(g["general_prior"])):
        rows, n = (int(v) for v in g["general_shape"][i])
        m = g["general_in"][g["general_in_off"][i]:g["general_in_off"][i + 1]].reshape(rows, n)
        p = float(g["general_prior"][i])
        exp = g["general_out"][g["general_out_off"][i]:g["general_out_off"][i + 1]]
        out.append((m, None if p < 0 else p, int(g["general_tri"][i]), exp))
    return out


def mixture_two_cases():
    """(matrix, prior, expected 2x2) of mixture_two.cpp:9."""
    g = load_mixture()
    out, off = [], 0
    for i, rows in enumerate(g["two_rows"]):
        m = g["two_in"][off:off + 2 * rows].reshape(rows, 2)
        off += 2 * rows
        out.append((m, float(g["two_prior"][i]), g["two_out"][i].reshape(2, 2)))
    return out
