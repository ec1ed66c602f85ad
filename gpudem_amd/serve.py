"""Serving: a request loop over one ``Batch`` (continuous batching).

Requests arrive on a wall-clock schedule.  Each loop iteration enqueues a
chunk of batch steps, and while the device runs them, admits the requests
that have arrived, in arrival order, as scene slots and capacity (the
batch's HBM budget) allow.  The head of the queue waits when it does not fit;
first-fit placement reuses the ranges of evicted scenes.  After the chunk, finished
scenes are downloaded and evicted.  Every admitted scene keeps its state on
the device and steps together with all others in the batched kernels, so a
request that arrives while others run overlaps with them on the GPU.
"""
import collections
import time

import numpy as np

from . import engine
from .engine import Batch


class Request:
    __slots__ = ("rid", "arrival", "spec", "nsteps", "sid", "admitted", "done", "result", "stats", "error",
                 "remaining", "snap", "preemptions", "start_step", "error_bits")

    def __init__(self, rid, arrival, spec, nsteps):
        self.rid = rid
        self.arrival = arrival
        self.spec = spec
        self.nsteps = nsteps
        self.sid = None
        self.admitted = None
        self.done = None
        self.result = None
        self.stats = None
        self.error = None
        self.remaining = nsteps  # steps still to run
        self.snap = None         # state and histories of an evicted run, resumed on the next admission
        self.preemptions = 0
        self.start_step = None   # batch step of the current admission
        self.error_bits = 0      # error bits of earlier residencies


class Server:
    """Continuous-batching server of DEM scenes on one device.

    chunk: batch steps per loop iteration (the scheduling quantum);
    keep_results: download every finished scene's final state;
    preempt_after: when the head of the queue has waited this long (seconds)
    and does not fit, the running request with the most steps left -- at
    least preempt_min_steps, and more than the head has left -- is evicted
    with a snapshot of its state and contact histories and queued again right
    behind the head; it resumes from the snapshot, bit for bit as if it had
    never left (None: never preempt).  A preemption always hands the GPU to a
    request with fewer steps left, so two requests never evict each other in
    turn.
    """

    def __init__(self, ctx, precision, hbm_budget=8 << 30, max_scenes=256, chunk=200, graph_steps=100,
                 nb_cap=32, wall_cap=16, hist_cap=16, whist_cap=8, keep_results=True, preempt_after=None,
                 preempt_min_steps=1000, batch=None):
        self.ctx = ctx
        self.chunk = chunk
        self.keep_results = keep_results
        self.preempt_after = preempt_after
        self.preempt_min_steps = preempt_min_steps
        self.preemptions = 0
        # `batch`: an object with Batch's interface to serve on instead (the
        # scheduling tests use a host-only stand-in)
        self.batch = batch if batch is not None else Batch(
            ctx, precision=precision, max_scenes=max_scenes, hbm_budget=hbm_budget, nb_cap=nb_cap,
            wall_cap=wall_cap, hist_cap=hist_cap, whist_cap=whist_cap, graph_steps=graph_steps)

    def _admit(self, req, now):
        """True if admitted, False if it must wait for capacity; raises
        GdemError when the request can never be admitted (larger than the
        batch, or invalid)."""
        s = req.spec
        m, k, e = engine.material(s.r, s.density, s.young, s.cor)
        if req.snap is None:
            sid = self.batch.admit(s.x, s.r, m, k, e, s.verts, s.params, req.remaining, v=s.v)
        else:
            sn = req.snap
            sid = self.batch.admit(sn["x"], s.r, m, k, e, s.verts, s.params, req.remaining, v=sn["v"],
                                   angv=sn["angv"], active=sn["active"])
            if sid is not None:
                self.batch.set_history(sid, sn)
                req.snap = None
        if sid is None:
            return False
        req.sid = sid
        req.start_step = self.batch.info()["step"]
        if req.admitted is None:
            req.admitted = now
        return True

    def _preempt(self, running, end_of, head, now):
        """Evicts the running request with the most steps left, keeping a
        snapshot; returns it, or None when none qualifies."""
        if self.preempt_after is None or now - head.arrival < self.preempt_after:
            return None
        step = self.batch.info()["step"]
        left = {sid: end_of[sid] - step for sid in running}
        sid = max(left, key=left.get, default=None)
        if sid is None or left[sid] < self.preempt_min_steps or left[sid] <= head.remaining:
            return None
        req = running.pop(sid)
        end_of.pop(sid)
        self.ctx.sync()
        req.error_bits |= self.batch.status(sid, check=False)["errors"]
        req.snap = self.batch.snapshot(sid)
        req.remaining -= self.batch.info()["step"] - req.start_step
        self.batch.evict(sid)
        req.sid = None
        req.preemptions += 1
        self.preemptions += 1
        return req

    def run(self, stream):
        """stream: iterable of (arrival_s, spec, nsteps).  Returns the list of
        finished Requests (arrival, admitted and done times in seconds since
        the start of the run)."""
        pending = collections.deque(Request(i, a, s, n) for i, (a, s, n) in
                                    enumerate(sorted(stream, key=lambda t: t[0])))
        queue = collections.deque()
        running = {}
        finished = []
        end_of = {}
        t0 = time.perf_counter()
        b = self.batch
        while pending or queue or running:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            if running:
                # run up to the earliest completion (at least a few steps)
                step = b.info()["step"]
                k = min(self.chunk, max(1, min(end_of[sid] - step for sid in running)))
                b.step(k)
            # admissions overlap the chunk on the device; they start after it
            preempted = False  # at most one preemption per chunk
            while queue:
                req = queue[0]
                try:
                    admitted = self._admit(req, time.perf_counter() - t0)
                except engine.GdemError as err:
                    # can never run here (larger than the whole batch, or invalid): fail it
                    queue.popleft()
                    req.error = str(err)
                    req.done = time.perf_counter() - t0
                    finished.append(req)
                    continue
                if not admitted:
                    if not running:
                        # refused with nothing running: no eviction will make room
                        queue.popleft()
                        req.error = "refused by an empty batch"
                        req.done = time.perf_counter() - t0
                        finished.append(req)
                        continue
                    victim = None if preempted else self._preempt(running, end_of, req, time.perf_counter() - t0)
                    if victim is None:
                        break
                    preempted = True
                    queue.insert(1, victim)  # resumes right after the head
                    continue
                queue.popleft()
                running[req.sid] = req
                end_of[req.sid] = b.info()["step"] + req.remaining
            if not running:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            self.ctx.sync()
            done_now = time.perf_counter() - t0
            for sid in b.poll():
                req = running.pop(sid)
                end_of.pop(sid)
                req.stats = b.status(sid, check=False)
                # steps of earlier residencies (before a preemption) count too
                req.stats["steps_done"] += req.nsteps - req.remaining
                req.stats["errors"] |= req.error_bits
                if self.keep_results:
                    req.result = b.download(sid)
                req.done = done_now
                b.evict(sid)
                finished.append(req)
        self.makespan = time.perf_counter() - t0
        return finished

    def close(self):
        self.batch.close()


def summarize(finished, makespan):
    """Throughput and latency percentiles of a finished run (failed requests
    are counted apart)."""
    failed = [r for r in finished if r.error is not None]
    finished = [r for r in finished if r.error is None]
    lat = np.array([r.done - r.arrival for r in finished])
    wait = np.array([r.admitted - r.arrival for r in finished])
    work = float(sum(r.spec.n * r.nsteps for r in finished))
    return dict(requests=len(finished), failed=len(failed), particle_steps=work, makespan_s=makespan,
                value=work / makespan if makespan > 0 else 0.0,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), queue_wait_p50_s=float(np.percentile(wait, 50)),
                queue_wait_p99_s=float(np.percentile(wait, 99)),
                errors=int(sum(1 for r in finished if r.stats and r.stats["errors"])),
                preemptions=int(sum(r.preemptions for r in finished)))
