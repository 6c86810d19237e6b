"""Per-class launch choices picked from timed sweeps on the caller's own batch.

The engine runs one persistent kernel per rows-per-lane class (1..16), so one
timed pass of a candidate setting gives every class's kernel time at once
(per-class HIP events, ``vd_class_timing``); a sweep costs candidates x reps
passes, not candidates x classes.  Choices (``vd_tuning`` in include/vdhmm.h):

* forward: pair units, haplotype prefix sharing, or prefix sharing with the
  shared-suffix continuation; anti-diagonals per main-loop iteration (2 or 4,
  pair units) x persistent grid per CU (the occupancy limit, or 2 or 3 blocks
  per CU), the per-class picks kept only when the whole forward beats the
  default choices; before the per-class sweep, the haplotypes per read unit under
  sharing (4, 8 or 16), timed for the whole forward with the class kernels
  on two streams and on one; after it, the stream layout
  (``vd_tuning.class_overlap``) with the final choices. Two streams fill each
  kernel's tail, but where one class holds most of the work they share the
  CUs with it for its whole run, and their time varies more: they are kept
  only when more than 2% faster;
* forward: then whether every class runs pair units packed across regions
  (``vd_tuning.fwd_pack``, no prefix sharing; kept when more than 2% faster
  end to end): for batches of regions with one or a few pairs, where
  region-bound units leave most of a wave's lanes idle;
* forward, precision "mixed": when the FP32 pass flags at least a quarter of
  the pairs, whether the active classes go straight to the FP64 rescue
  (``vd_tuning.fwd_route64``; kept when more than 2% faster, and the routed
  pairs then get their FP64 values);
* adjoint: register block of 2 steps (two waves per SIMD where that fits) or
  4 steps.

None of them changes the arithmetic of a pair except the suffix
continuation (which sums the suffix columns through an adjoint and agrees
with full sweeps to FP32 rounding) and the FP64 routing (routed pairs get
their FP64 values), so results are otherwise identical for every choice;
only the time differs.  ``tune_forward`` / ``tune_grad`` apply
the per-class best and return a report; ``save`` / ``load`` keep a choice
(every field of ``vd_tuning``, the stream layout included) in a JSON file
keyed by workload and mode, and ``load`` gives a field the entry lacks its
default.  The choices are the library's, for every later call of the
process: tune on the batch that will run, or load its saved choice.
"""
from __future__ import annotations

import ctypes
import json
from pathlib import Path

import numpy as np

from . import _lib

# (unroll, blocks per CU (0 = occupancy limit), sharing: 0 pair units, 1 prefix, 2 prefix +
# suffix for every read unit, 3 suffix kernel for regions with a shared suffix and the
# prefix-only kernel for the others); the shared kernels have one unroll.  The first is the
# default.
FWD_CANDIDATES = [(2, 0, 2), (2, 0, 1), (2, 0, 3), (2, 0, 0), (4, 0, 0), (2, 3, 2), (2, 3, 1), (2, 3, 0), (4, 3, 0),
                  (2, 2, 2), (2, 2, 1), (2, 2, 0)]
CHUNK_CANDIDATES = (4, 8, 16)
GRAD_CANDIDATES = [(0, 0), (1, 0)]                           # (variant, blocks per CU)


def get() -> _lib.VdTuning:
    t = _lib.VdTuning()
    _lib.check(_lib.lib().vd_tuning_get(ctypes.byref(t)), "vd_tuning_get")
    return t


def set_tuning(t: _lib.VdTuning):
    _lib.check(_lib.lib().vd_tuning_set(ctypes.byref(t)), "vd_tuning_set")


def class_ms():
    f = (ctypes.c_double * 17)()
    g = (ctypes.c_double * 17)()
    _lib.lib().vd_class_ms(f, g)
    return np.array(f[:]), np.array(g[:])


def _sweep(run, candidates, apply, which: int, reps: int):
    """Per candidate: the minimum over reps of each class's kernel time."""
    L = _lib.lib()
    base = get()
    times = {}
    L.vd_class_timing(1)
    try:
        for cand in candidates:
            t = _lib.VdTuning.from_buffer_copy(base)
            for r in range(1, 17):
                apply(t, r, cand)
            set_tuning(t)
            best = np.full(17, np.inf)
            for _ in range(reps):
                run()
                best = np.minimum(best, class_ms()[which])
            times[cand] = best
    finally:
        L.vd_class_timing(0)
        set_tuning(base)
    return times


def _pick(times, candidates, active):
    choice = {}
    for r in range(1, 17):
        if not active[r]:
            continue
        choice[r] = min(candidates, key=lambda c: times[c][r])
    return choice


def _apply_fwd(t, r, cand):
    t.fwd_unroll[r], t.fwd_blocks_per_cu[r], t.fwd_share[r] = cand


def _apply_grad(t, r, cand):
    t.grad_variant[r], t.grad_blocks_per_cu[r] = cand


def tune_forward(dev_batch, mode=0, precision="mixed", reps=3, candidates=FWD_CANDIDATES):
    """Sweep the forward choices on a DeviceBatch and apply the best per class.

    Each candidate's class times are the minimum over `reps` passes (with 2,
    run-to-run noise of a few percent flipped close classes between pair units
    and prefix sharing and cost up to 3% of the pass).  The per-class picks are
    then checked as a whole against the default choices, whose kernels overlap
    on two streams the same way, and kept only when the whole forward is faster."""
    def run():
        dev_batch.forward(mode=mode, precision=precision, stats=True)

    t0 = get()
    t0.fwd_pack = 0  # the per-class sweep times region-bound units
    set_tuning(t0)
    run()  # warm: buffers and tables
    lib = _lib.lib()
    pair_ms = {}
    base = get()
    for overlap in (1, 0):  # whole-forward time of each stream layout x read-unit size (sharing on everywhere)
        for chunk in CHUNK_CANDIDATES:
            t = _lib.VdTuning.from_buffer_copy(base)
            t.share_chunk = chunk
            t.class_overlap = overlap
            for r in range(1, 17):
                t.fwd_share[r] = 2
            set_tuning(t)
            best = np.inf
            for _ in range(reps):
                best = min(best, dev_batch.forward(mode=mode, precision=precision, stats=True)["fwd32_ms"])
            pair_ms[(overlap, chunk)] = best
    overlap, chunk = min(pair_ms, key=pair_ms.get)
    chunk_ms = {c: pair_ms[(overlap, c)] for c in CHUNK_CANDIDATES}
    t = _lib.VdTuning.from_buffer_copy(base)
    t.share_chunk = chunk
    set_tuning(t)
    times = _sweep(run, candidates, _apply_fwd, 0, reps)
    default = times[candidates[0]]
    active = default > 0.02  # classes with work (an empty class kernel exits at once)
    choice = _pick(times, candidates, active)
    t = get()
    for r, c in choice.items():
        _apply_fwd(t, r, c)
    set_tuning(t)
    tuned = sum(times[c][r] for r, c in choice.items())
    # the picks as a whole against the default choices (whole-forward time)
    def whole():
        return min(dev_batch.forward(mode=mode, precision=precision, stats=True)["fwd32_ms"] for _ in range(reps))
    picked_ms = whole()
    t_def = _lib.VdTuning.from_buffer_copy(t)
    for r in range(1, 17):
        _apply_fwd(t_def, r, candidates[0])
    set_tuning(t_def)
    default_whole_ms = whole()
    if picked_ms <= default_whole_ms:
        set_tuning(t)
    else:
        choice = {r: candidates[0] for r in choice}
        tuned = sum(times[candidates[0]][r] for r in choice)
    # the stream layout with the final choices: two streams are kept only when
    # they beat one by more than 2%, since their time varies more from run to
    # run (it depends on how the two kernels' blocks interleave)
    overlap_ms = {}
    for o in (1, 0):
        lib.vd_class_overlap(o)
        overlap_ms[o] = min(dev_batch.forward(mode=mode, precision=precision, stats=True)["fwd32_ms"]
                            for _ in range(max(reps, 2)))
    overlap = 1 if overlap_ms[1] < 0.98 * overlap_ms[0] else 0
    lib.vd_class_overlap(overlap)   # kept in vd_tuning.class_overlap with the other choices
    report = {"kind": "forward", "choice": {r: list(c) for r, c in choice.items()},
              "share_chunk": int(get().share_chunk), "chunk_ms": {k: round(float(v), 3) for k, v in chunk_ms.items()},
              "class_overlap": int(overlap), "overlap_ms": {k: round(float(v), 3) for k, v in overlap_ms.items()},
              "default_ms": round(float(np.sum(default[active])), 3), "tuned_ms": round(float(tuned), 3),
              "whole_ms": {"picked": round(float(picked_ms), 3), "default": round(float(default_whole_ms), 3)}}
    report["pack"] = _tune_pack(dev_batch, mode, precision, reps)
    if precision == "mixed":
        report["route64"] = _tune_route(dev_batch, mode, active, reps)
    return report


def _tune_pack(dev_batch, mode, precision, reps):
    """Time the whole forward with the tuned choices and with every class on
    pair units packed across regions (vd_tuning.fwd_pack); keep the packing
    when it is more than 2% faster."""
    def total():
        return min(dev_batch.forward(mode=mode, precision=precision, stats=True)["device_ms"]
                   for _ in range(max(reps, 2)))
    base = get()
    base_ms = total()
    t = _lib.VdTuning.from_buffer_copy(base)
    t.fwd_pack = 1
    set_tuning(t)
    packed_ms = total()
    keep = packed_ms < 0.98 * base_ms
    if not keep:
        set_tuning(base)
    return {"fwd_pack": int(keep), "tuned_ms": round(float(base_ms), 3), "packed_ms": round(float(packed_ms), 3)}


def _tune_route(dev_batch, mode, active, reps, min_flagged=0.25):
    """When the FP32 pass flags at least `min_flagged` of the batch's pairs
    for the FP64 rescue, time the batch with its active classes sent straight
    to FP64 (vd_tuning.fwd_route64) and keep that when it is more than 2%
    faster end to end."""
    def total():
        best, st = np.inf, None
        for _ in range(max(reps, 2)):
            st = dev_batch.forward(mode=mode, precision="mixed", stats=True)
            best = min(best, st["device_ms"])
        return best, st
    base = get()
    base_ms, st = total()
    frac = st["flagged"] / max(st["pairs"], 1)
    out = {"flagged_frac": round(float(frac), 4), "classes": [], "mixed_ms": round(float(base_ms), 3)}
    if frac < min_flagged:
        return out
    t = _lib.VdTuning.from_buffer_copy(base)
    classes = [r for r in range(1, 17) if active[r]]
    for r in classes:
        t.fwd_route64[r] = 1
    set_tuning(t)
    routed_ms, _ = total()
    out["routed_ms"] = round(float(routed_ms), 3)
    if routed_ms < 0.98 * base_ms:
        out["classes"] = classes
    else:
        set_tuning(base)
    return out


def tune_grad(dev_batch, mode=0, reps=1, candidates=GRAD_CANDIDATES):
    """Sweep the adjoint choices on a DeviceBatch and apply the best per class."""
    def run():
        dev_batch.grad(mode=mode, stats=True)

    run()
    times = _sweep(run, candidates, _apply_grad, 1, reps)
    default = times[candidates[0]]
    active = default > 0.02
    choice = _pick(times, candidates, active)
    t = get()
    for r, c in choice.items():
        _apply_grad(t, r, c)
    set_tuning(t)
    tuned = sum(times[c][r] for r, c in choice.items())
    return {"kind": "adjoint", "choice": {r: list(c) for r, c in choice.items()},
            "default_ms": round(float(np.sum(default[active])), 3), "tuned_ms": round(float(tuned), 3)}


def defaults() -> _lib.VdTuning:
    """The library's launch choices before any tuning (include/vdhmm.h)."""
    t = _lib.VdTuning()
    for r in range(17):
        t.fwd_unroll[r], t.fwd_blocks_per_cu[r], t.fwd_share[r] = FWD_CANDIDATES[0]
        t.grad_variant[r], t.grad_blocks_per_cu[r] = GRAD_CANDIDATES[0]
        t.fwd_route64[r] = 0
    t.share_chunk = CHUNK_CANDIDATES[0]
    t.class_overlap = 1
    t.fwd_pack = 0
    return t


def save(path, key: str):
    p = Path(path)
    data = json.loads(p.read_text()) if p.exists() else {}
    data[key] = get().as_dict()
    p.write_text(json.dumps(data, indent=1, sort_keys=True))


def load(path, key: str) -> bool:
    """Apply the choice kept under ``key``; a field the entry lacks takes its
    default, so no choice of an earlier tuning leaks into the loaded one."""
    p = Path(path)
    if not p.exists():
        return False
    data = json.loads(p.read_text())
    if key not in data:
        return False
    t = defaults()
    for name, _ in t._fields_:
        if name not in data[key]:
            continue
        if name in _lib.VdTuning.SCALARS:
            setattr(t, name, int(data[key][name]))
        else:
            arr = getattr(t, name)
            for r, v in enumerate(data[key][name][:17]):
                arr[r] = int(v)
    set_tuning(t)
    return True
