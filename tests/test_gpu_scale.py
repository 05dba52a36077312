"""The batched (C4) and served (C5) forms at the sizes their workloads name, on the GPU.

C4: 256 meshes of the 16.8M-node workload's family, sizes log-uniform in [5k, 500k] nodes (seeded),
dam breaks and still-water basins over their own bumps, FCT + alpha, all in one batch for 10
steps: every instance equals its mesh run alone on a fresh solver (state, time, step count), and
three members of about 5k, 50k and 500k nodes equal the oracle after 3 steps.

C5: a Poisson stream of 24 requests drawn by serve_request_specs (lognormal sizes, median 100k,
one request of 1.14M nodes, 5 dry-bed dam breaks with Manning friction = 21%), served by the
continuous-batching loop into a batch whose capacity comes from a memory budget too small to hold
every request at once (so requests wait for rows), admitted from their element lists (set-up and
layout on the GPU).  Every served result (state, time, steps, error) equals the same request run
alone on a fresh solver from the host assembly; depths stay >= 0; the two smallest requests equal
the oracle.  Final times are 1 to 20 wave periods of 1 s (the workload's 20 s periods scaled down
so the alone runs fit a test).

Each instance follows SOURCES/main.f90:61-91 (the loop runs while time < tfinal).
"""
import numpy as np
import pytest

from conftest import rel_diff, same_bits

pytestmark = pytest.mark.gpu

C5_SEED = 20261236     # 24 requests: max 1,135,575 nodes, 5 dry-bed dam breaks, median 109,507


def oracle_of(oracle, p):
    pr = p.params
    visc = {"galerkin": oracle.GALERKIN, "viscous": oracle.VISCOUS, "entropy_visc": oracle.ENTROPY_VISC,
            "fct": oracle.FCT}[pr["viscosity"]]
    return oracle.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=pr["gravity"], htiny=pr["htiny"],
                         cfl=pr["cfl"], max_water_h=pr["max_water_h"], viscosity_type=visc,
                         viscous_type=pr["viscous_type"], alpha_limit=pr["alpha_limit"],
                         type_test=pr["type_test"], mannings=pr["mannings"])


def oracle_run(oracle, p, nsteps):
    hd, uxd, uyd = p.dirichlet()
    kind, amp, period, level = p.bc
    orc = oracle_of(oracle, p)
    u, dts, t, err = orc.run(p.u0, nsteps, t0=0.0, bc_kind=kind, amp=amp, period=period, level=level,
                             nodes_h=hd, nodes_ux=uxd, nodes_uy=uyd)
    orc.close()
    return u, t, err


def agrees(a, b, friction):
    """Bitwise, or within 1e-11 relative L1 and Linf where Manning friction (pow) is on."""
    if friction:
        l1, linf = rel_diff(a, b)
        return l1 <= 1e-11 and linf <= 1e-11
    return same_bits(a, b)


def test_c4_256_instances_equal_runs_alone_and_the_oracle(oracle):
    from swe_amd import problems
    from swe_amd.batch import INSTANCE_DONE, Batch
    specs = problems.batch_family_specs(256, 5000, 500000)
    ps = problems.build_family(specs, workers=16)
    nodes = sum(p.np for p in ps)
    print("C4: %d meshes, %d nodes (%d..%d)" % (len(ps), nodes, min(p.np for p in ps), max(p.np for p in ps)),
          flush=True)
    n = 10
    rows = sum((p.np + 63) // 64 * 64 for p in ps)
    entries = int(sum(p.nnz for p in ps) * 1.35)
    bnd = sum(sum(x.size for x in p.dirichlet()) for p in ps) + 64
    with Batch(rows, entries, bnd, len(ps)) as b:
        ids = [b.add_problem(p, max_steps=n) for p in ps]
        assert None not in ids
        b.step(n + 1)
        res = {k: (b.status(k), b.get_state(k)) for k in ids}
    print("C4: batch done; comparing each instance with its mesh alone", flush=True)
    for k, p in zip(ids, ps):
        (st, tb, steps, err), u = res[k]
        assert st == INSTANCE_DONE and steps == n and err == 0, (p.name, st, steps, err)
        with p.solver() as s:
            s.set_state(p.u0)
            t, _ = s.run(n, time=0.0)
            assert t == tb and same_bits(s.get_state(), u), p.name
        assert u[:, 0].min() >= 0.0
    # three members against the oracle, 3 steps
    order = np.argsort([p.np for p in ps])
    picks = [order[0], order[int(np.searchsorted(np.sort([p.np for p in ps]), 50000))], order[-1]]
    for q in picks:
        p = ps[q]
        print("C4: oracle on %s (%d nodes), 3 steps" % (p.name, p.np), flush=True)
        ref, t, err = oracle_run(oracle, p, 3)
        assert err == 0
        with p.solver() as s:
            s.set_state(p.u0)
            tg, _ = s.run(3, time=0.0)
            assert tg == t and same_bits(s.get_state(), ref), p.name


def test_c5_served_stream_equals_requests_alone(oracle):
    from swe_amd import problems
    from swe_amd.batch import Batch
    from swe_amd.serve import BatchEngine, Request, Server
    specs = problems.serve_request_specs(24, 50.0, seed=C5_SEED, period=1.0)
    ps = problems.build_family(specs, workers=16)
    sizes = [p.np for p in ps]
    dry = sum(1 for s in specs if s["kind"] == "drydam")
    assert max(sizes) >= 1_000_000 and dry / len(specs) >= 0.2
    print("C5: %d requests, %d nodes (max %d, median %d), %d dry-bed" % (len(ps), sum(sizes), max(sizes),
                                                                        int(np.median(sizes)), dry), flush=True)
    # a budget for about 1.6M nodes: the 1.14M-node request fits, every request at once does not
    budget = Batch.bytes_for(1_600_000, 9 * 1_600_000, 32_000, 64)
    reqs = [Request(rid=i, problem=p, tfinal=s["tfinal"], arrival=s["arrival"]) for i, (p, s) in
            enumerate(zip(ps, specs))]
    with Batch.for_budget(budget, max_instances=64) as b:
        assert b.caps.max_nodes < sum((x + 63) // 64 * 64 for x in sizes)
        srv = Server(BatchEngine(b, device_setup=True), chunk=8)
        done = srv.run(reqs)
    print("C5: served, at most %d resident" % srv.max_resident, flush=True)
    assert srv.max_resident < len(reqs)
    first_done = min(r.finished for r in done)
    assert any(r.admitted > first_done for r in done)          # someone waited for rows to free up
    for r in done:
        p = r.problem
        assert r.error == 0 and r.time >= r.tfinal and r.steps > 0, (r.rid, r.error, r.time, r.tfinal)
        assert r.state[:, 0].min() >= 0.0
        with p.solver() as s:
            s.set_state(p.u0)
            t1, _ = s.run(r.steps - 1, time=0.0) if r.steps > 1 else (0.0, None)
            assert t1 < r.tfinal                                 # the loop took exactly these steps
            t, _ = s.run(1)
            assert t == r.time and same_bits(s.get_state(), r.state), (r.rid, p.name)
    for r in sorted(done, key=lambda r: r.problem.np)[:2]:
        p = r.problem
        print("C5: oracle on request %d (%s, %d nodes, %d steps)" % (r.rid, p.name, p.np, r.steps), flush=True)
        ref, t, err = oracle_run(oracle, p, r.steps)
        assert err == 0 and t == r.time
        assert agrees(r.state, ref, p.params["mannings"] > 0), (r.rid, p.name)


def test_c4_island_members_stop_where_the_reference_stops(oracle):
    """The C4 members whose bumps emerge as dry islands (a quarter of the family), 1000 steps in one
    batch: the instances that the reference would stop with 'Min h<0' (check_hmin) fail at a step
    of their own while the others go on; each failed instance reports the same step, time and error
    when run alone, and for the smallest of them the oracle runs without error up to the step before
    (the same bits) and reports the negative depth at that step."""
    from swe_amd import problems
    from swe_amd.batch import INSTANCE_DONE, INSTANCE_FAILED, Batch
    from swe_amd.solver import SWE_ENEG_H
    specs = [s for s in problems.batch_family_specs(256, 5000, 500000) if s["zmax"] is None]
    ps = problems.build_family(specs, workers=16)
    n = 1000
    print("C4 islands: %d meshes, %d nodes" % (len(ps), sum(p.np for p in ps)), flush=True)
    rows = sum((p.np + 63) // 64 * 64 for p in ps)
    entries = int(sum(p.nnz for p in ps) * 1.35)
    bnd = sum(sum(x.size for x in p.dirichlet()) for p in ps) + 64
    with Batch(rows, entries, bnd, len(ps)) as b:
        ids = [b.add_problem(p, max_steps=n, device_setup=True) for p in ps]
        b.step(n + 1)
        stat = {k: b.status(k) for k in ids}
    failed = [(p, stat[k]) for k, p in zip(ids, ps) if stat[k][0] == INSTANCE_FAILED]
    print("C4 islands: %d failed at steps %s" % (len(failed), sorted(st[2] for _, st in failed)), flush=True)
    assert all(stat[k][0] in (INSTANCE_DONE, INSTANCE_FAILED) for k in ids)
    assert all(stat[k][2] == n for k in ids if stat[k][0] == INSTANCE_DONE)
    assert failed and all(st[3] == SWE_ENEG_H and st[2] <= n for _, st in failed)

    def alone(p, steps):
        with Batch((p.np + 63) // 64 * 64, int(p.nnz * 1.4) + 4096, sum(x.size for x in p.dirichlet()) + 64, 1) as b1:
            k = b1.add_problem(p, max_steps=steps, device_setup=True)
            b1.step(steps + 1)
            return b1.status(k), b1.get_state(k)
    for p, st in failed:
        assert alone(p, n)[0] == st, p.name
    p, st = min(failed, key=lambda x: x[0].np)
    before, u = alone(p, st[2] - 1)
    assert before[0] == INSTANCE_DONE and before[2] == st[2] - 1
    print("C4 islands: oracle on %s (%d nodes), %d steps" % (p.name, p.np, st[2]), flush=True)
    hd, uxd, uyd = p.dirichlet()
    kind, amp, period, level = p.bc
    orc = oracle_of(oracle, p)
    try:
        kw = dict(bc_kind=kind, amp=amp, period=period, level=level, nodes_h=hd, nodes_ux=uxd, nodes_uy=uyd)
        ref, _, t, err = orc.run(p.u0, st[2] - 1, t0=0.0, **kw)
        assert err == 0 and t == before[1] and same_bits(u, ref)
        _, _, _, err = orc.run(ref, 1, t0=t, **kw)
        assert err == oracle.ENEG_H
    finally:
        orc.close()
