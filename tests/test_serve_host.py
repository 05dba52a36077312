"""The serving loop's host logic on the CPU, over a simulated engine: every request completes, the
capacity is never exceeded, admission is FIFO and never precedes arrival, and each request takes
the steps its final time requires (the same as when served alone)."""
import math

import pytest

from swe_amd.serve import Request, SchemeEngine, Server, run_plain, scheme_of, summarize


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-4          # every look at the clock costs a little time
        return self.t

    def sleep(self, s):
        self.t += max(0.0, s)


class FakeProblem:
    def __init__(self, n, dt, params=None):
        self.np = n
        self.dt = dt
        self.params = params or {}


class FakeEngine:
    """Instances of `np` rows in a pool of `capacity` rows; each step advances an instance's time by
    its dt while time < tfinal; a step costs time proportional to the resident rows.  step()
    returns a ticket whose record lists the instances finished when that chunk ended; an
    instance admitted after a ticket is not in its record."""

    def __init__(self, capacity, clock):
        self.capacity = capacity
        self.clock = clock
        self.inst = {}
        self.next = 0
        self.max_rows = 0
        self.order = []
        self.tickets = []
        self.reported = set()
        self.drained = 0          # reads of the newest ticket while instances were still running

    def add(self, r):
        used = sum(i["np"] for i in self.inst.values())
        if used + r.problem.np > self.capacity:
            return None
        k = self.next
        self.next += 1
        self.inst[k] = dict(np=r.problem.np, dt=r.problem.dt, tfinal=r.tfinal, time=0.0, steps=0)
        self.order.append(r.rid)
        self.max_rows = max(self.max_rows, used + r.problem.np)
        return k

    def step(self, n):
        for _ in range(n):
            for i in self.inst.values():
                if i["time"] < i["tfinal"]:
                    i["time"] = i["time"] + i["dt"]
                    i["steps"] += 1
            self.clock.t += 1e-6 * sum(i["np"] for i in self.inst.values())
        self.tickets.append({k for k, i in self.inst.items() if not i["time"] < i["tfinal"]})
        return len(self.tickets) - 1

    def finished(self, ticket):
        if ticket == len(self.tickets) - 1 and any(i["time"] < i["tfinal"] for i in self.inst.values()):
            self.drained += 1
        out = sorted(k for k in self.tickets[ticket] if k in self.inst and k not in self.reported)
        self.reported.update(out)
        return out

    def collect(self, k):
        assert k in self.reported                     # only reported instances are collected
        i = self.inst.pop(k)
        return None, i["time"], i["steps"], 0


def stream(n=40, seed=3):
    import numpy as np
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(0.01, n))
    return [Request(rid=q, problem=FakeProblem(int(rng.integers(50, 400)), float(rng.uniform(0.05, 0.2))),
                    tfinal=float(rng.uniform(1.0, 20.0)), arrival=float(arr[q])) for q in range(n)]


@pytest.mark.parametrize("chunk", [1, 8, 32])
def test_server_completes_every_request_within_capacity(chunk):
    clock = FakeClock()
    eng = FakeEngine(1000, clock)
    reqs = Server(eng, chunk=chunk, clock=clock, sleep=clock.sleep).run(stream())
    assert all(r.finished is not None for r in reqs)
    assert eng.max_rows <= 1000
    assert eng.order == sorted(eng.order)                       # FIFO admission
    assert eng.drained == 0                                     # a chunk is always queued ahead
    for r in reqs:
        assert r.admitted >= r.arrival and r.finished >= r.admitted
        # DO WHILE (time < tfinal): the first step count whose time reaches tfinal
        t, n = 0.0, 0
        while t < r.tfinal:
            t = t + r.problem.dt
            n += 1
        assert r.steps == n and r.time == t
    s = summarize(reqs)
    assert s["requests"] == 40 and s["p99_s"] >= s["p50_s"] > 0 and s["errors"] == 0


def test_plain_loop_takes_the_same_steps_and_is_serial():
    clock = FakeClock()
    served = Server(FakeEngine(1000, clock), chunk=8, clock=clock, sleep=clock.sleep).run(stream())
    clock2 = FakeClock()
    eng = FakeEngine(400, clock2)
    plain = run_plain(eng, stream(), chunk=8, clock=clock2, sleep=clock2.sleep)
    assert [r.steps for r in plain] == [r.steps for r in served]
    assert [r.time for r in plain] == [r.time for r in served]
    assert eng.max_rows <= 400
    for a, b in zip(plain, plain[1:]):
        assert b.admitted >= a.finished                          # one at a time


def test_request_larger_than_the_batch_is_refused():
    clock = FakeClock()
    r = [Request(rid=0, problem=FakeProblem(5000, 0.1), tfinal=1.0, arrival=0.0)]
    with pytest.raises(RuntimeError):
        Server(FakeEngine(1000, clock), clock=clock, sleep=clock.sleep).run(r)


def test_idle_server_waits_for_arrivals():
    clock = FakeClock()
    reqs = [Request(rid=q, problem=FakeProblem(100, 0.5), tfinal=1.0, arrival=10.0 * q) for q in range(3)]
    out = Server(FakeEngine(1000, clock), clock=clock, sleep=clock.sleep).run(reqs)
    assert [math.floor(r.admitted / 10.0) for r in out] == [0, 1, 2]


def test_several_schemes_are_served_side_by_side():
    """One engine per scheme behind SchemeEngine: each request lands in its scheme's engine, within
    that engine's capacity, and takes the steps its final time requires, as when served alone."""
    clock = FakeClock()
    schemes = [dict(viscosity="fct", alpha_limit=True), dict(viscosity="viscous", alpha_limit=True),
               dict(viscosity="viscous", viscous_type=1), dict(viscosity="galerkin", alpha_limit=True)]
    reqs = stream(n=48, seed=5)
    for q, r in enumerate(reqs):
        r.problem.params = schemes[q % 4]
    made = {}

    def make(sch):
        made[sch] = FakeEngine(500, clock)
        return made[sch]
    eng = SchemeEngine(make)
    out = Server(eng, chunk=4, clock=clock, sleep=clock.sleep).run(reqs)
    assert len(made) == 4 and scheme_of(schemes[3]) == ("galerkin", 2, False, True, False)
    for sch, e in made.items():
        assert e.max_rows <= 500 and not e.inst
        assert e.order == sorted(e.order) and {scheme_of(reqs[q].problem.params) for q in e.order} == {sch}
    for r in out:
        t, n = 0.0, 0
        while t < r.tfinal:
            t = t + r.problem.dt
            n += 1
        assert r.steps == n and r.time == t and r.finished >= r.admitted >= r.arrival
