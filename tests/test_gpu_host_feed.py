"""The host feed with the device scorer (needs an MI355X): pooled coroutine
requests and evidence actors whose requests are scored in the parent by the
batched engine give the same likelihoods as scoring each request directly
(bit for bit: a pair's FP64 result does not depend on its place in a launch)."""
import numpy as np
import pytest

from vardepth_amd import accelerator
from vardepth_amd.batch_coroutines import score_generators
from vardepth_amd.gpu_evidence_pipeline import run_tasks

pytestmark = pytest.mark.gpu


def _pairs(task, count):
    rng = np.random.default_rng(1000 + task)
    hap = "".join(rng.choice(list("ACGT"), int(rng.integers(120, 300))))
    out = []
    for _ in range(count):
        n = int(rng.integers(20, 120))
        s0 = int(rng.integers(0, len(hap) - n))
        read = hap[s0:s0 + n]
        qual = bytes(rng.integers(10, 41, n).astype(np.uint8))
        out.append((read, qual, hap))
    return out


def evidence(task):
    """Two rounds of requests, the second in the other mode, sized by the first's answer."""
    first = _pairs(task, 3 + task % 4)
    v1 = yield bool(task % 2), first
    second = _pairs(task + 100, 1 + int(np.argmax(v1)))
    v2 = yield not bool(task % 2), second
    return [float(x) for x in v1], [float(x) for x in v2]


def test_score_generators_with_device_scorer():
    def gen(task):
        jobs = _pairs(task, 2 + task % 5)
        values = yield jobs
        more = yield jobs[:1]
        return list(values) + list(more)

    calls = []

    def score(jobs):
        calls.append(len(jobs))
        return accelerator.score_pairs(jobs)

    out = score_generators([gen(t) for t in range(30)], score, max_pairs=16)
    assert max(calls) <= 16
    for t in range(30):
        jobs = _pairs(t, 2 + t % 5)
        direct = accelerator.score_pairs(jobs)
        np.testing.assert_array_equal(np.array(out[t]), np.concatenate([direct, direct[:1]]))


def test_run_tasks_with_device_scorer():
    tasks = list(range(12))
    out = run_tasks(tasks, None, None, workers=2, slots_per_worker=3, max_pairs=10, factory=evidence)
    for t, (v1, v2) in zip(tasks, out):
        first = _pairs(t, 3 + t % 4)
        d1 = accelerator.score_pairs(first, anchored=bool(t % 2))
        second = _pairs(t + 100, 1 + int(np.argmax(d1)))
        d2 = accelerator.score_pairs(second, anchored=not bool(t % 2))
        np.testing.assert_array_equal(np.array(v1), d1)
        np.testing.assert_array_equal(np.array(v2), d2)
