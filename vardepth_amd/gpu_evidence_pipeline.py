"""Evidence actors in CPU worker processes, likelihoods on the GPU in the parent.

Drop-in for the reference's ``gpu_evidence_pipeline.run_tasks``
(caller/native/caller/gpu_evidence_pipeline.py:74-132, actor loop :43-71).
Each task runs in a spawned worker as a generator made by ``factory(task)``;
the generator yields ``(anchored, requests)`` -- a mode flag and a list of
(read, quality, haplotype) jobs -- and receives the likelihood array.  Only
the requests and the arrays cross process boundaries; the parent pools the
suspended requests of all workers and scores them with one call per anchored
mode and flush (``scorer(requests, anchored)``, by default
:func:`vardepth_amd.accelerator.score_pairs`, one persistent device launch).

Bounds: ``workers`` processes, ``slots_per_worker`` live tasks per worker
(least-loaded assignment), flushes of up to ``max_pairs`` jobs, taken when
every live task is waiting for likelihoods or when ``max_pairs`` jobs are
pending.  Results come back in task order.  A worker exception, a dead
worker or a scorer of the wrong shape raises in the parent after every child
is stopped.

The reference's default factory is its own evidence chain
(``full_steps``: partial anchoring, graph proposals, whole-read evidence over
BAM input), which is outside this library: ``factory`` is required whenever
there are tasks.  ``arguments`` / ``reference``, when ``arguments`` is not
None, go to ``initialize(arguments, reference)`` in every worker before its
first task (the reference calls ``evidence_workers.initialize``).
"""
from __future__ import annotations

import itertools
import multiprocessing as mp
import queue
import traceback

import numpy as np


def _actor(inbox, outbox, arguments, reference, factory, initialize):
    import os
    # workers only suspend likelihood requests; the parent owns the device
    os.environ['VARDEPTH_PAIRHMM_BACKEND'] = 'cuda'
    live = {}
    ident = None
    try:
        if arguments is not None:
            initialize(arguments, reference)
        while True:
            message = inbox.get()
            if message is None:
                return
            kind, ident, payload = message
            if kind == 'start':
                if ident in live:
                    raise ValueError('Duplicate active evidence task')
                gen = live[ident] = factory(payload)
                send = None
                fresh = True
            elif kind == 'resume':
                gen = live[ident]
                send = payload
                fresh = False
            else:
                raise ValueError('Unknown evidence actor command')
            try:
                anchored, jobs = next(gen) if fresh else gen.send(send)
            except StopIteration as stop:
                del live[ident]
                outbox.put(('done', ident, stop.value))
            else:
                outbox.put(('request', ident, bool(anchored), jobs))
    except BaseException:
        outbox.put(('error', ident, traceback.format_exc()))
    finally:
        for gen in live.values():
            gen.close()


class _Pipeline:
    def __init__(self, tasks, arguments, reference, workers, slots_per_worker, max_pairs, scorer, factory,
                 initialize):
        self.tasks = tasks
        self.workers = workers
        self.window = workers * slots_per_worker
        self.max_pairs = max_pairs
        self.scorer = scorer
        ctx = mp.get_context('spawn')
        self.outbox = ctx.Queue(maxsize=2 * self.window)
        self.inboxes = [ctx.Queue(maxsize=2 * slots_per_worker) for _ in range(workers)]
        self.procs = [ctx.Process(target=_actor, args=(self.inboxes[w], self.outbox, arguments, reference,
                                                       factory, initialize)) for w in range(workers)]
        self.owner = {}             # live task -> worker
        self.load = [0] * workers
        self.results = []
        self.pending = []           # (task, anchored, jobs) waiting for likelihoods
        self.busy = 0               # tasks computing in a worker (sent, no reply yet)
        self.exhausted = False

    def admit(self):
        while not self.exhausted and len(self.owner) < self.window:
            try:
                task = next(self.tasks)
            except StopIteration:
                self.exhausted = True
                return
            w = min(range(self.workers), key=self.load.__getitem__)
            ident = len(self.results)
            self.results.append(None)
            self.owner[ident] = w
            self.load[w] += 1
            self.inboxes[w].put(('start', ident, task))
            self.busy += 1

    def flush(self):
        mode = self.pending[0][1]
        take, keep, count = [], [], 0
        for item in self.pending:
            if item[1] == mode and (not take or count + len(item[2]) <= self.max_pairs):
                take.append(item)
                count += len(item[2])
            else:
                keep.append(item)
        self.pending = keep
        values = np.asarray(self.scorer([job for _, _, jobs in take for job in jobs], mode))
        if values.ndim != 1 or len(values) != count:
            raise ValueError('Likelihood scorer returned the wrong shape')
        at = 0
        for ident, _, jobs in take:
            self.inboxes[self.owner[ident]].put(('resume', ident, values[at:at + len(jobs)]))
            at += len(jobs)
            self.busy += 1

    def run(self):
        for p in self.procs:
            p.start()
        self.admit()
        while self.owner:
            if self.pending and (self.busy == 0 or sum(len(x[2]) for x in self.pending) >= self.max_pairs):
                self.flush()
                continue
            try:
                kind, ident, *payload = self.outbox.get(timeout=2)
            except queue.Empty:
                if any(not p.is_alive() for p in self.procs):
                    raise RuntimeError('Evidence worker exited before completing its tasks')
                continue
            if kind == 'error':
                raise RuntimeError('Evidence worker failed:\n' + payload[0])
            if ident not in self.owner:
                raise RuntimeError('Unowned evidence response')
            self.busy -= 1
            if kind == 'request':
                self.pending.append((ident, payload[0], payload[1]))
            elif kind == 'done':
                self.results[ident] = payload[0]
                self.load[self.owner.pop(ident)] -= 1
                self.admit()
            else:
                raise RuntimeError('Unknown evidence response')
        if self.pending or self.busy:
            raise RuntimeError('Evidence accounting mismatch')
        return self.results

    def close(self, clean):
        if clean:
            for q in self.inboxes:
                q.put(None)
            for p in self.procs:
                p.join(timeout=10)
        for p in self.procs:
            if p.is_alive():
                p.terminate()
                p.join(timeout=2)
        for q in self.inboxes + [self.outbox]:
            if not clean:
                q.cancel_join_thread()
            q.close()
            if clean:
                q.join_thread()


def run_tasks(tasks, arguments, reference, *, workers=4, slots_per_worker=8, max_pairs=32768, scorer=None,
              factory=None, initialize=None):
    if workers < 1 or slots_per_worker < 1 or max_pairs < 1:
        raise ValueError('Invalid evidence pipeline bounds')
    tasks = iter(tasks)
    try:
        first = next(tasks)
    except StopIteration:
        return []
    if factory is None:
        raise ValueError('run_tasks needs an evidence factory (the reference evidence chain is not part of '
                         'this library)')
    if arguments is not None and initialize is None:
        raise ValueError('run_tasks needs initialize(arguments, reference) when arguments are given')
    if scorer is None:
        from .accelerator import backend, score_pairs
        choice = backend()
        scorer = lambda requests, anchored: score_pairs(requests, anchored=anchored, selection=choice)  # noqa: E731
    pipe = _Pipeline(itertools.chain([first], tasks), arguments, reference, workers, slots_per_worker, max_pairs,
                     scorer, factory, initialize)
    clean = False
    try:
        out = pipe.run()
        clean = True
        return out
    finally:
        pipe.close(clean)
