"""A queue of atmosphere states solved on one GPU: the host side of the batched form (the role of the reference's
state queue and persistent GPU service, `src/sosfuse2/jobs/pipeline.py:409-678` run_queue and
`src/sosfuse2/jobs/pipeline_execution.py`, and of its chunk loop's next-chunk preparation on a second thread,
`src/sosfuse2/jobs/run_state.py:833-869`).

`StateQueue.submit(state, wavelengths, elevations)` returns a `Job` at once and may be called from any thread. One
worker thread owns the GPU: it takes the queued states, packs them into chunks of up to `lanes` lanes (first-fit
decreasing on the wavelength count, batch.pack), and solves each chunk as one source iteration (each chunk's
evaluation a captured HIP graph) followed by its table query. While the GPU works on a chunk, a second thread
prepares the next chunk's host inputs (the lanes' optics and their solver context). A finished job holds its table
[elevations, view elevations, azimuths, W, 3] (Stokes I, Q, U) and the chunk's solve record, and with `out_dir` set
each state's table is written as `<out_dir>/<name>/table.npz` in the reference's layout (api.save_table). A chunk
that fails (no convergence, non-finite source) fails its jobs with that error and the queue goes on with the next.
"""

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

from . import api, atmosphere, batch


class Job:
    """One queued state: wait() returns (table, info) or raises the chunk's error."""

    def __init__(self, item, name):
        self.item = item
        self.name = name
        self.table = None
        self.info = None
        self.error = None
        self.path = None
        self._done = threading.Event()

    def done(self):
        return self._done.is_set()

    def wait(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError(f'E_QUEUE_TIMEOUT: {self.name} not finished within {timeout} s')
        if self.error is not None:
            raise self.error
        return self.table, self.info

    def _finish(self, table=None, info=None, error=None):
        self.table, self.info, self.error = table, info, error
        self._done.set()


class StateQueue:
    """States of one grid and profile shape (the scope's) solved in packed chunks by one GPU worker thread.

    solve_chunk(scope, items, el, az, context) -> [(table, info)] is the GPU step (the default runs the source
    iteration and the table query of batch.solve_chunk); a test may replace it."""

    def __init__(self, scope, el, az, tolerance=1e-5, max_eval=60, lanes=32, out_dir=None, linger=0.02,
                 x0=None, solve_chunk=None):
        self.scope = scope
        self.el, self.az = np.asarray(el, float), np.asarray(az, float)
        self.tolerance, self.max_eval, self.lanes = tolerance, max_eval, lanes
        self.out_dir = None if out_dir is None else Path(out_dir)
        self.linger = linger  # seconds the worker waits for more states before it packs what it holds
        self.x0 = x0
        self.solve_chunk = solve_chunk or self._solve_chunk
        self._pending = []
        self._cv = threading.Condition()
        self._closed = False
        self._count = 0
        self.stats = dict(chunks=0, states=0, failed=0, seconds=0.0)
        self._prep = ThreadPoolExecutor(max_workers=1)
        self._worker = threading.Thread(target=self._run, name='sos-state-queue', daemon=True)
        self._worker.start()

    # ---------------------------------------------------------------------------------------------- clients
    def submit(self, state, wavelengths, elevations, name=None):
        """Queue one state; raises E_BATCH_PROFILE at once for a state the scope's geometry cannot serve."""
        atmosphere.check_profile_shapes(state, self.scope.state)
        wl = np.atleast_1d(np.asarray(wavelengths, float))
        if not 1 <= len(wl) <= self.lanes:
            raise ValueError(f'E_WAVELENGTHS: {len(wl)} wavelengths for a {self.lanes}-lane queue')
        with self._cv:
            if self._closed:
                raise RuntimeError('E_QUEUE_CLOSED: the queue takes no more states')
            self._count += 1
            job = Job(batch.Item(state, wl, list(elevations)), name or f'{state.name}_{self._count}')
            self._pending.append(job)
            self._cv.notify()
        return job

    def close(self, wait=True):
        """Take no more states; with wait, return once every queued state is finished."""
        with self._cv:
            self._closed = True
            self._cv.notify()
        if wait:
            self._worker.join()
            self._prep.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ----------------------------------------------------------------------------------------------- worker
    def _take(self):
        """Wait for states, linger briefly for more, and return one chunk's jobs (None when closed and empty)."""
        with self._cv:
            while not self._pending and not self._closed:
                self._cv.wait()
            if not self._pending:
                return None
            if not self._closed and self.linger > 0:
                self._cv.wait(self.linger)
            jobs = self._pending
            chunks = batch.pack([j.item for j in jobs], self.lanes)
            take = [jobs[i] for i in chunks[0].members]
            self._pending = [j for j in jobs if j not in take]
            return take

    def _prepare(self, jobs):
        """Host inputs of a chunk (on the preparation thread): the packed lanes' optics and the solver context."""
        optics = atmosphere.concat_optics([atmosphere.lane_optics(j.item.state, j.item.wavelengths) for j in jobs])
        return jobs, self.scope.context(optics)

    def _solve_chunk(self, scope, items, el, az, context):
        return batch.solve_chunk(scope, items, el, az, context, tolerance=self.tolerance, max_eval=self.max_eval,
                                 x0=self.x0)

    def _run(self):
        nxt = None
        while True:
            if nxt is None:
                jobs = self._take()
                if jobs is None:
                    return
                nxt = self._prep.submit(self._prepare, jobs)
            t = time.perf_counter()
            try:
                jobs, ctx = nxt.result()
            except Exception as exc:  # a failed preparation fails its jobs
                for j in jobs:
                    j._finish(error=exc)
                nxt = None
                continue
            # the next chunk's host preparation overlaps this chunk's GPU work
            more = self._take_nowait()
            nxt = self._prep.submit(self._prepare, more) if more else None
            try:
                results = self.solve_chunk(self.scope, [j.item for j in jobs], self.el, self.az, ctx)
                for j, (table, info) in zip(jobs, results):
                    if self.out_dir is not None:
                        d = self.out_dir / j.name
                        d.mkdir(parents=True, exist_ok=True)
                        j.path = api.save_table(d / 'table.npz', table, j.item.wavelengths, j.item.elevations,
                                                self.el, self.az)
                    j._finish(table, info)
                self.stats['states'] += len(jobs)
            except Exception as exc:
                for j in jobs:
                    j._finish(error=exc)
                self.stats['failed'] += len(jobs)
            self.stats['chunks'] += 1
            self.stats['seconds'] += time.perf_counter() - t

    def _take_nowait(self):
        with self._cv:
            if not self._pending:
                return None
            chunks = batch.pack([j.item for j in self._pending], self.lanes)
            take = [self._pending[i] for i in chunks[0].members]
            self._pending = [j for j in self._pending if j not in take]
            return take
