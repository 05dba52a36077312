"""Serving: shallow-water solves of different sizes and final times, admitted as they arrive into
one device-resident batch and advanced together (continuous batching), with time-to-solution
recorded per request.

:class:`Server` runs a pipelined loop that never lets the GPU drain while requests are resident:
it queues the next chunk of steps (which ends with a copy of the instance table, its ticket),
then waits only for the previous chunk's table copy, reads the results of the instances that copy
shows finished (on the copy stream, while the queued chunk runs), frees their rows, and admits
waiting requests (FIFO) into the free rows, against the batch's capacity, which is set from a
memory budget.  A request joins the steps queued after its admission and leaves when its time
reaches its final time (the loop condition of SOURCES/main.f90:61), so its result is the one it
would get alone.  :func:`run_plain` serves the same stream one request at a time on the same
engine: the plain per-instance loop the server is measured against.

The engine interface (add / step / finished / collect) is what the loop needs from a batch:
``step(n)`` queues n steps and returns a ticket, ``finished(ticket)`` lists the instances that
ticket's table copy shows done or failed, ``collect(k)`` returns (state, time, steps, error) and
frees the instance.  :class:`BatchEngine` binds it to :class:`swe_amd.batch.Batch`.  A batch runs
one scheme (viscosity type, viscous type, alpha limiter, roundoff fix, SGN model): a stream that
mixes schemes is served by :class:`SchemeEngine`, one batch per scheme on its own stream, all
advanced by the same loop so that their steps overlap on the device.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

MAX_STEPS = 2 ** 62


@dataclasses.dataclass
class Request:
    rid: int
    problem: object              # swe_amd.problems.Problem
    tfinal: float
    arrival: float               # seconds after the start of the stream
    admitted: float | None = None
    finished: float | None = None
    time: float = 0.0
    steps: int = 0
    error: int = 0
    state: np.ndarray | None = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def node_stages(self):
        return self.problem.np * 3 * self.steps


class BatchEngine:
    """The serving loop's view of a :class:`swe_amd.batch.Batch`."""

    def __init__(self, batch, device_setup=False):
        self.batch = batch
        self.device_setup = device_setup     # admit from the element list: set-up and layout on the GPU

    def add(self, req: Request):
        return self.batch.add_problem(req.problem, tfinal=req.tfinal, max_steps=MAX_STEPS,
                                      device_setup=self.device_setup)

    def step(self, n):
        return self.batch.step_async(n)

    def finished(self, ticket):
        return self.batch.finished(ticket)

    def collect(self, k):
        st, u, t, steps, err = self.batch.result(k)
        self.batch.remove(k)
        return u, t, steps, err

    def close(self):
        self.batch.close()


def scheme_of(params):
    """What the instances of one batch share (swe_batch_add refuses the others): viscosity type,
    viscous type, alpha limiter (never with galerkin), roundoff fix, and the SGN model (type_test 13)."""
    visc = params.get("viscosity", "fct")
    return (visc, int(params.get("viscous_type", 2)), bool(params.get("alpha_limit", False)) and visc != "galerkin",
            bool(params.get("roundoff", True)), int(params.get("type_test", 4)) == 13)


class SchemeEngine:
    """The engine interface over one engine (batch) per scheme.  ``make_engine(scheme)`` creates the
    engine of a scheme the first time a request of that scheme arrives (a :class:`BatchEngine` over
    a batch whose capacity is that scheme's share of the memory budget); instance ids are
    (scheme, id) pairs and a ticket holds one ticket per engine that has instances."""

    def __init__(self, make_engine):
        self.make_engine = make_engine
        self.engines = {}
        self.live = collections.Counter()

    def add(self, req: Request):
        sch = scheme_of(req.problem.params)
        eng = self.engines.get(sch)
        if eng is None:
            eng = self.engines[sch] = self.make_engine(sch)
        k = eng.add(req)
        if k is None:
            return None
        self.live[sch] += 1
        return (sch, k)

    def step(self, n):
        return {sch: eng.step(n) for sch, eng in self.engines.items() if self.live[sch]}

    def finished(self, ticket):
        return [(sch, k) for sch, t in ticket.items() for k in self.engines[sch].finished(t)]

    def collect(self, key):
        sch, k = key
        self.live[sch] -= 1
        return self.engines[sch].collect(k)

    def close(self):
        for eng in self.engines.values():
            if hasattr(eng, "close"):
                eng.close()
        self.engines.clear()


class Server:
    """Continuous batching: requests join and leave a running batch between chunks of steps, with
    one chunk always queued ahead of the one the host is reading."""

    def __init__(self, engine, chunk=16, clock=time.perf_counter, sleep=time.sleep):
        self.engine = engine
        self.chunk = int(chunk)
        self.clock = clock
        self.sleep = sleep
        self.max_resident = 0

    def run(self, requests):
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        pending = collections.deque(reqs)
        waiting = collections.deque()
        active = {}
        prev = None
        t0 = self.clock()
        while pending or waiting or active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # queue the next chunk before reading the last one, so the GPU stays busy meanwhile
            ticket = self.engine.step(self.chunk) if active else None
            if prev is not None:
                for k in self.engine.finished(prev):
                    r = active.pop(k)
                    r.state, r.time, r.steps, r.error = self.engine.collect(k)
                    r.finished = self.clock() - t0
            while waiting:
                r = waiting[0]
                k = self.engine.add(r)
                if k is None:
                    if not active:
                        raise RuntimeError("request %d (%d nodes) does not fit an empty batch" % (r.rid, r.problem.np))
                    break
                waiting.popleft()
                active[k] = r
                r.admitted = self.clock() - t0
            prev = ticket
            self.max_resident = max(self.max_resident, len(active))
            if not active and not waiting and pending:
                self.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
        return reqs


def run_plain(engine, requests, chunk=16, clock=time.perf_counter, sleep=time.sleep):
    """The same stream served one request at a time, in arrival order: queue a chunk, wait for
    it, look whether the request is done."""
    reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
    t0 = clock()
    for r in reqs:
        now = clock() - t0
        if r.arrival > now:
            sleep(r.arrival - now)
        r.admitted = clock() - t0
        k = engine.add(r)
        if k is None:
            raise RuntimeError("request %d (%d nodes) does not fit the engine" % (r.rid, r.problem.np))
        while k not in engine.finished(engine.step(chunk)):
            pass
        r.state, r.time, r.steps, r.error = engine.collect(k)
        r.finished = clock() - t0
    return reqs


def summarize(reqs):
    """p50/p99/mean time-to-solution (s), aggregate node-stage updates/s over the makespan."""
    lat = np.array([r.latency for r in reqs])
    work = float(sum(r.node_stages for r in reqs))
    span = max(r.finished for r in reqs) - min(r.arrival for r in reqs)
    return {"requests": len(reqs), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": float(span), "node_stages": work,
            "node_stage_per_s": work / span if span > 0 else float("nan"),
            "errors": int(sum(1 for r in reqs if r.error))}
