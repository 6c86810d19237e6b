"""Host feed of the engine: pooled likelihood coroutines and the evidence-actor
pipeline (CPU; the scorer is a plain function here, the device scorer is
exercised in test_gpu_dropin.py)."""
import multiprocessing as mp

import numpy as np
import pytest

from vardepth_amd.batch_coroutines import score_generators
from vardepth_amd.gpu_evidence_pipeline import run_tasks


# ---- score_generators (batch_coroutines.py:4-32) -------------------------------

def test_followup_requests_are_rebatched_within_bound():
    calls = []

    def worker(i):
        a = yield [i]
        b = yield [i] * 7
        return float(a[0]), float(sum(b))

    def score(jobs):
        calls.append(len(jobs))
        return np.asarray(jobs, dtype=float)

    out = score_generators((worker(i) for i in range(20)), score, max_pairs=10)
    assert out == [(float(i), float(i * 7)) for i in range(20)]
    assert max(calls) <= 10
    assert sum(calls) == 20 * 8


def test_oversized_request_is_scored_alone_and_order_kept():
    calls = []

    def worker(i, n):
        v = yield list(range(n))
        return i, float(np.sum(v))

    def score(jobs):
        calls.append(len(jobs))
        return np.asarray(jobs, dtype=float)

    sizes = [3, 25, 2, 2, 9, 0]
    gens = [worker(i, n) for i, n in enumerate(sizes)]
    out = score_generators(gens, score, max_pairs=8)
    assert out == [(i, float(sum(range(n)))) for i, n in enumerate(sizes)]
    # 3 | 25 alone | 2+2 | 9 alone | the empty request: a pooled flush never exceeds 8
    assert calls == [3, 25, 4, 9, 0]


def test_generators_without_requests_and_bounds():
    def done(i):
        return i
        yield  # noqa: unreachable -- makes this a generator

    assert score_generators([done(1), done(2)], lambda jobs: jobs) == [1, 2]
    with pytest.raises(ValueError, match='max_pairs must be positive'):
        score_generators([], lambda jobs: jobs, max_pairs=0)


def test_bad_scorer_length_fails():
    def worker():
        yield [1, 2]

    with pytest.raises(ValueError, match='wrong number of results'):
        score_generators([worker()], lambda jobs: [])


# ---- run_tasks (gpu_evidence_pipeline.py:74-132) --------------------------------

def two_rounds(task):
    values = yield bool(task % 2), [task] * 2
    first = float(np.sum(values))
    values = yield not bool(task % 2), [task + 1] * (task % 5 + 1)
    return task, first + float(np.sum(values))


def fails_on_three(task):
    if task == 3:
        raise ValueError('planted worker failure')
    yield False, [task]
    return task


def scale_by_mode(jobs, anchored):
    return np.asarray(jobs, dtype=float) * (2 if anchored else 3)


def add_offset(arguments, reference):
    import os
    os.environ['VD_TEST_OFFSET'] = str(arguments + reference)


def offset_task(task):
    import os
    v = yield False, [task]
    return float(v[0]) + float(os.environ['VD_TEST_OFFSET'])


def test_pipeline_order_modes_bounds_and_followups():
    tasks = list(range(20, -1, -1))
    seen = []

    def score(jobs, anchored):
        seen.append((len(jobs), anchored))
        return scale_by_mode(jobs, anchored)

    out = run_tasks(tasks, None, None, workers=3, slots_per_worker=3, max_pairs=4, scorer=score, factory=two_rounds)
    expected = [(x, 2 * x * (2 if x % 2 else 3) + (x % 5 + 1) * (x + 1) * (3 if x % 2 else 2)) for x in tasks]
    assert out == expected
    assert {a for _, a in seen} == {True, False}
    assert all(n <= 5 for n, _ in seen)      # one request may exceed max_pairs (up to 5 jobs), never a pooled flush


def test_pipeline_worker_initializer():
    out = run_tasks([1, 2, 3], 10, 5, workers=2, slots_per_worker=1, scorer=scale_by_mode, factory=offset_task,
                    initialize=add_offset)
    assert out == [3.0 + 15, 6.0 + 15, 9.0 + 15]


def test_pipeline_worker_error_stops_children():
    before = {p.pid for p in mp.active_children()}
    with pytest.raises(RuntimeError, match='planted worker failure'):
        run_tasks(range(12), None, None, workers=2, slots_per_worker=2, scorer=scale_by_mode, factory=fails_on_three)
    assert {p.pid for p in mp.active_children()} == before


def test_pipeline_bad_scorer_stops_children():
    before = {p.pid for p in mp.active_children()}
    with pytest.raises(ValueError, match='wrong shape'):
        run_tasks(range(12), None, None, workers=2, slots_per_worker=2,
                  scorer=lambda jobs, a: np.zeros(len(jobs) + 1), factory=two_rounds)
    assert {p.pid for p in mp.active_children()} == before


def test_pipeline_bounds_and_empty():
    assert run_tasks([], None, None, scorer=scale_by_mode, factory=two_rounds) == []
    for kw in (dict(workers=0), dict(slots_per_worker=0), dict(max_pairs=0)):
        with pytest.raises(ValueError):
            run_tasks([], None, None, **kw)
    with pytest.raises(ValueError, match='evidence factory'):
        run_tasks([1], None, None, scorer=scale_by_mode)
