"""Batched execution of likelihood coroutines (the host feed of the engine).

Drop-in for the reference's ``batch_coroutines.score_generators``
(caller/native/caller/batch_coroutines.py:4-32).  Each generator yields a
list of (read, quality, haplotype) jobs, receives their likelihoods, and may
yield again; its return value is its result.  Requests from many generators
are pooled so that one call of ``score`` -- one persistent device launch in
:func:`vardepth_amd.accelerator.score_pairs` -- covers up to ``max_pairs``
jobs; a single request larger than that is scored on its own, never split.
Results come back in generator order.
"""
from __future__ import annotations

from collections import deque


def score_generators(generators, score, max_pairs=32768):
    if max_pairs < 1:
        raise ValueError('max_pairs must be positive')
    results = []
    waiting = deque()          # (index, generator, request) whose request is not scored yet
    queued = 0                 # jobs held in `waiting` since the last drain

    def step(index, generator, values=None, first=False):
        """Advance one generator; the next (index, generator, request) or None when it finished."""
        try:
            request = next(generator) if first else generator.send(values)
        except StopIteration as stop:
            results[index] = stop.value
            return None
        return index, generator, request

    def drain():
        nonlocal queued
        while waiting:
            take = [waiting.popleft()]
            count = len(take[0][2])
            while waiting and count + len(waiting[0][2]) <= max_pairs:
                take.append(waiting.popleft())
                count += len(take[-1][2])
            jobs = [job for _, _, request in take for job in request]
            values = score(jobs)
            if len(values) != len(jobs):
                raise ValueError('Likelihood scorer returned the wrong number of results')
            at = 0
            for index, generator, request in take:
                follow = step(index, generator, values[at:at + len(request)])
                at += len(request)
                if follow is not None:
                    waiting.append(follow)     # follow-up requests join this drain's later rounds
        queued = 0

    for index, generator in enumerate(generators):
        results.append(None)
        item = step(index, generator, first=True)
        if item is None:
            continue
        if queued and queued + len(item[2]) > max_pairs:
            drain()
        waiting.append(item)
        queued += len(item[2])
        if queued >= max_pairs:
            drain()
    drain()
    return results
