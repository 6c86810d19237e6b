"""Evidence actors in CPU worker processes, likelihoods on the GPU in the parent.

Drop-in for ``gpu_evidence_pipeline.run_tasks``
(caller/native/caller/gpu_evidence_pipeline.py:74-132): same arguments,
results, bounds and error messages.

Contract.  Each task runs in a spawned worker as the generator
``factory(task)``.  The generator yields ``(anchored, requests)`` -- a mode
flag and a list of (read, quality, haplotype) jobs -- and is sent back the
likelihood array; its return value is the task's result.  Only requests,
arrays and results cross process boundaries; generators (and whatever
evidence objects they hold) stay in their worker.  The parent pools the
suspended requests of all workers and scores them with one
``scorer(requests, anchored)`` call per flush (by default
:func:`vardepth_amd.accelerator.score_pairs`: one persistent device launch).

Scheduling, as the reference defines it:

* at most ``workers * slots_per_worker`` tasks are live; a new task goes to
  the worker with the fewest live tasks, and results come back in task order;
* a flush happens when some request waits and either no task is computing
  in a worker, or the waiting requests hold ``max_pairs`` jobs or more;
* a flush serves one mode, that of the oldest waiting request: waiting
  requests of that mode are taken oldest first while they fit within
  ``max_pairs`` (the first one is always taken, however large);
* a worker exception, a dead worker, a reply for an unknown task or a scorer
  result of the wrong shape raises in the parent, after every child is
  stopped.

Design.  Each worker talks to the parent over its own duplex pipe.  The
parent's event loop waits on every pipe and every worker's process sentinel
at once (``multiprocessing.connection.wait``), so a reply is handled as it
arrives and a worker that dies is seen at once.  Messages to a worker go
through a per-worker sender thread, so the event loop never blocks on a full
pipe while that worker is itself blocked sending a request.

The reference's default factory is its own evidence chain (``full_steps``:
partial anchoring, graph proposals, whole-read evidence over BAM input),
which is outside this library: ``factory`` is required whenever there are
tasks.  With ``arguments`` not None, every worker first calls
``initialize(arguments, reference)`` (the reference calls
``evidence_workers.initialize``).
"""
from __future__ import annotations

import multiprocessing as mp
import queue
import threading
import traceback
from multiprocessing.connection import wait

import numpy as np

_STOP = None        # the message that ends a worker
_QUIT = object()    # ends a sender thread (never sent)


def _serve(conn, arguments, reference, factory, initialize):
    """Worker process: drive the generators of the tasks the parent starts here."""
    import os
    # a worker only suspends likelihood requests; the parent owns the device
    os.environ['VARDEPTH_PAIRHMM_BACKEND'] = 'cuda'
    running = {}
    current = None
    try:
        if arguments is not None:
            initialize(arguments, reference)
        while True:
            msg = conn.recv()
            if msg is _STOP:
                return
            verb, current, payload = msg
            if verb == 'start':
                if current in running:
                    raise ValueError('Duplicate active evidence task')
                running[current] = factory(payload)
                advance = lambda gen: next(gen)                     # noqa: E731
            elif verb == 'resume':
                advance = lambda gen, v=payload: gen.send(v)        # noqa: E731
            else:
                raise ValueError('Unknown evidence actor command')
            try:
                anchored, jobs = advance(running[current])
            except StopIteration as finished:
                del running[current]
                conn.send(('done', current, finished.value))
            else:
                conn.send(('request', current, bool(anchored), jobs))
    except BaseException:
        try:
            conn.send(('error', current, traceback.format_exc()))
        except (OSError, ValueError):
            pass
    finally:
        for gen in running.values():
            gen.close()
        conn.close()


class _Worker:
    """Parent-side handle: process, pipe, and the thread that feeds the pipe."""

    def __init__(self, ctx, arguments, reference, factory, initialize):
        self.conn, child = ctx.Pipe(duplex=True)
        self.proc = ctx.Process(target=_serve, args=(child, arguments, reference, factory, initialize))
        self.proc.start()
        child.close()
        self.live = 0
        self.outgoing = queue.Queue()
        self.sender = threading.Thread(target=self._pump, daemon=True)
        self.sender.start()

    def _pump(self):
        while True:
            msg = self.outgoing.get()
            if msg is _QUIT:
                return
            try:
                self.conn.send(msg)
            except (OSError, ValueError):
                return          # the worker is gone; the event loop reports it

    def post(self, msg):
        self.outgoing.put(msg)

    def shut(self, clean: bool):
        """Ask the worker to exit (clean) and stop the sender thread."""
        if clean:
            self.post(_STOP)
        self.outgoing.put(_QUIT)


class _Scheduler:
    def __init__(self, tasks, workers, slots, max_pairs, scorer):
        self.tasks = tasks
        self.window = len(workers) * slots
        self.max_pairs = max_pairs
        self.scorer = scorer
        self.workers = workers
        self.owner = {}            # live task -> its worker
        self.results = []
        self.waiting = []          # [task, anchored, jobs], oldest first
        self.computing = 0         # tasks inside a worker's generator (no reply yet)
        self.exhausted = False

    def admit(self):
        while not self.exhausted and len(self.owner) < self.window:
            try:
                task = next(self.tasks)
            except StopIteration:
                self.exhausted = True
                return
            w = min(self.workers, key=lambda x: x.live)
            ident = len(self.results)
            self.results.append(None)
            self.owner[ident] = w
            w.live += 1
            w.post(('start', ident, task))
            self.computing += 1

    def due(self) -> bool:
        return bool(self.waiting) and (self.computing == 0 or
                                       sum(len(x[2]) for x in self.waiting) >= self.max_pairs)

    def flush(self):
        mode = self.waiting[0][1]
        batch, rest, count = [], [], 0
        for item in self.waiting:
            if item[1] == mode and (not batch or count + len(item[2]) <= self.max_pairs):
                batch.append(item)
                count += len(item[2])
            else:
                rest.append(item)
        self.waiting = rest
        values = np.asarray(self.scorer([job for _, _, jobs in batch for job in jobs], mode))
        if values.ndim != 1 or len(values) != count:
            raise ValueError('Likelihood scorer returned the wrong shape')
        edges = np.cumsum([0] + [len(jobs) for _, _, jobs in batch])
        for (ident, _, _), lo, hi in zip(batch, edges[:-1], edges[1:]):
            self.owner[ident].post(('resume', ident, values[lo:hi]))
            self.computing += 1

    def handle(self, msg):
        verb, ident, *payload = msg
        if verb == 'error':
            raise RuntimeError('Evidence worker failed:\n' + payload[0])
        if ident not in self.owner:
            raise RuntimeError('Unowned evidence response')
        self.computing -= 1
        if verb == 'request':
            self.waiting.append([ident, payload[0], payload[1]])
        elif verb == 'done':
            self.results[ident] = payload[0]
            self.owner.pop(ident).live -= 1
            self.admit()
        else:
            raise RuntimeError('Unknown evidence response')

    def run(self):
        by_conn = {w.conn: w for w in self.workers}
        by_sentinel = {w.proc.sentinel: w for w in self.workers}
        self.admit()
        while self.owner:
            if self.due():
                self.flush()
                continue
            for ready in wait(list(by_conn) + list(by_sentinel)):
                if ready in by_conn:
                    try:
                        msg = ready.recv()
                    except (EOFError, OSError):
                        raise RuntimeError('Evidence worker exited before completing its tasks') from None
                    self.handle(msg)
                elif not any(c.poll() for c in by_conn):
                    # a worker ended and nothing it sent is left unread
                    raise RuntimeError('Evidence worker exited before completing its tasks')
        if self.waiting or self.computing:
            raise RuntimeError('Evidence accounting mismatch')
        return self.results


def _stop_all(workers, clean: bool):
    for w in workers:
        w.shut(clean)
    if clean:
        for w in workers:
            w.proc.join(timeout=10)
    for w in workers:
        if w.proc.is_alive():
            w.proc.terminate()
            w.proc.join(timeout=2)
    for w in workers:
        w.sender.join(timeout=2)
        w.conn.close()


def run_tasks(tasks, arguments, reference, *, workers=4, slots_per_worker=8, max_pairs=32768, scorer=None,
              factory=None, initialize=None):
    if workers < 1 or slots_per_worker < 1 or max_pairs < 1:
        raise ValueError('Invalid evidence pipeline bounds')
    tasks = iter(tasks)
    first = next(tasks, _STOP)
    if first is _STOP:
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

    def chained():
        yield first
        yield from tasks

    ctx = mp.get_context('spawn')
    pool = []
    clean = False
    try:
        for _ in range(workers):
            pool.append(_Worker(ctx, arguments, reference, factory, initialize))
        out = _Scheduler(chained(), pool, slots_per_worker, max_pairs, scorer).run()
        clean = True
        return out
    finally:
        _stop_all(pool, clean)
