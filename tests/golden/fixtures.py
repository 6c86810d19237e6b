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
