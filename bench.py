"""Benchmark of the MI355X explicit shallow-water step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload bumps|dam_break|paraboloid]

A step is one SSP-RK3 step (COMPUTE_DT + three Euler stages + bdy) over the whole mesh, with the
state resident in HBM.  The default workload is the 16.8M-node FCT + alpha problem (still water
over 64 Gaussian bumps with a sinusoidal inflow, 4097^2 jittered-lattice Delaunay nodes).
Prints one JSON line: node-stage updates/s, the HBM roofline of the dominant kernel (measured
with HIP events on the solver's stream), and the CPU checker's rate on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "shallow-water-eric_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
MESH_NAMES = {
    "bumps": "Delaunay triangulation of a seeded Poisson-disk point set (unstructured)",
    "bumps_lattice": "cell-split jittered lattice (structured topology)",
    "dam_break": "cell-split jittered lattice 2241 x 449",
    "paraboloid": "81 x 81 grid",
    "sgn": "the SGN canal [0,10]x[0,1], cell-split jittered lattice 6401 x 641",
}


def kernel_bytes_per_node(kernel, k, alpha=True, nc=3):
    """Algorithmic HBM bytes per node of one launch (one graph sweep, each array read once).
    G = 4 + 20k: row pointer, column index and c_ij (16 B) per entry.  nc = 5 (the SGN model): the
    stage pass also reads and writes h*eta, h*w (16 B each way) and reads them at U^n in stages 2
    and 3 (2/3 of 16 B on average)."""
    G = 4 + 20 * k
    if nc == 5 and kernel == "k_stage_lo":
        return kernel_bytes_per_node(kernel, k, alpha) + 16 + 16 + 16 * 2.0 / 3.0
    return {
        "k_fct_lo": G + 80,        # U,z,m,alpha in (48); U^L, res out (32)
        "k_fct_rminus": G + 72,    # U,z (32), res (8), U^L.h... (m, hL, res in; R out)
        "k_fct_apply": G + 112,    # U,z, res, R, U^L, U^n in; U out
        "k_stage_lo": 96 + 26.67 * k if not alpha else 124 + 30.67 * k,
        "k_compute_dt": G + 32,
        "k_alpha": 4 + 4 * k + 16,
    }[kernel]


def step_bytes_per_node_stage(k, fct=True, alpha=True, nc=3):
    """Compulsory HBM bytes per node-stage of the whole step, k = stored entries per row: the sum of
    the per-launch figures of kernel_bytes_per_node over one stage (FCT: three graph sweeps plus
    the alpha pass and a third of the dt pass), i.e. 324 + 70.67k with FCT and alpha; nc = 5 adds
    the SGN model's two components as in kernel_bytes_per_node."""
    if fct:
        return 324 + 70.67 * k
    extra = (16 + 16 + 16 * 2.0 / 3.0) if nc == 5 else 0.0
    return ((124 + 30.67 * k) if alpha else (96 + 26.67 * k)) + extra


def measured_traffic(kernel, workload, nodes):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE/WRITE_SIZE (calibrated
    against a plain device copy) for this workload and mesh size, from the committed summaries in
    profiles/, or None when no summary covers it."""
    pdir = os.path.join(REPO, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".txt"):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                parts = line.split()
                if (len(parts) == 5 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel
                        and parts[2] == workload and int(parts[3]) == nodes):
                    return float(parts[4])
    return None


def measured_valu(kernel, workload, nodes):
    """(VALU busy fraction, active-lane fraction) of `kernel` over a step's launches from the committed
    rocprofv3 VALU summary for this workload and mesh size (profiles/), or None."""
    pdir = os.path.join(REPO, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".txt"):
            continue
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                parts = line.split()
                if (len(parts) == 6 and parts[0] == "valu_per_launch" and parts[1] == kernel
                        and parts[2] == workload and int(parts[3]) == nodes):
                    return float(parts[4]), float(parts[5])
    return None


def build_problem(name, size=None, host_assembly=True):
    from swe_amd import problems
    if name == "bumps":
        return problems.bumps_unstructured(n_nodes=size or 16_785_409, host_assembly=host_assembly)
    if name == "bumps_lattice":
        return problems.bumps(n_side=size or 4097)
    if name == "dam_break":
        nx = size or 2241
        return problems.dam_break(nx=nx, ny=(nx - 1) // 5 + 1)
    if name == "paraboloid":
        return problems.paraboloid(n=size or 81)
    if name == "sgn":
        nx = size or 6401
        return problems.sgn_canal(nx=nx, ny=(nx - 1) // 10 + 1)
    raise SystemExit("unknown workload %s" % name)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """The cores this process may run on (the box's CPU share), at most 16."""
    try:
        cores = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cores = list(range(os.cpu_count() or 1))
    return cores[:16]


def reference_baseline(name, p, seconds, driver=None, all_cores=True):
    """The reference Fortran (oracle/_ref: the reference modules built with amdflang -O2, their
    SSP-RK3 loop of SOURCES/main.f90:61-92) on the mesh of problem `p`, timed by the driver's own
    clock around its time loop (set-up excluded).  driver = REF_DRIVER_HOIST is the build whose
    entropy_residual assigns its rescale array once instead of once per node (the same bits in O(np)
    rather than O(np^2) per stage).  One process pinned to one core with taskset: a short run gives
    the rate, a second run as many steps as fit in about `seconds`.  all_cores: then one process per
    core of this process's CPU share (at most 16), each running the same steps at once, their
    rates summed: the throughput of the whole host on independent meshes."""
    import shutil
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as O
    test = {"bumps": 4, "dam_break": 10, "paraboloid": 5, "sgn": 13}[name]
    cores = host_cores()
    pin = cores[0] if shutil.which("taskset") else None
    d = O.tmpdir()
    try:
        init = O.ref_case(d, p, test)
        nc = 5 if test == 13 else 3
        el0 = O.ref_time_loop(d, init, 1, driver=driver, pin=pin, out="one", nc=nc)[3]
        n = max(1, int(seconds / max(el0, 1e-6)))
        el = O.ref_time_loop(d, init, n, driver=driver, pin=pin, out="pinned", nc=nc)[3]
        res = {"value": p.np * 3 * n / el, "unit": "node-stage updates/s", "cores": 1, "kind": "reference",
               "cpu_model": cpu_model(), "pinned_core": pin,
               "sample": "%s, %d nodes, %d steps (%.1f s of time loop), the reference Fortran (amdflang -O2 build of "
                         "SOURCES/update.f90, fct.f90 and the RK3 loop%s), 1 thread%s"
                         % (p.name, p.np, n, el, ", rescale loop of update.f90:865-867 hoisted" if driver else "",
                            ", taskset -c %d" % pin if pin is not None else "")}
        if all_cores and len(cores) > 1:
            import subprocess as sp
            procs = []
            for q, c in enumerate(cores):
                od = os.path.join(d, "all%d" % q)
                os.makedirs(od, exist_ok=True)
                cmd = ["taskset", "-c", str(c)] if pin is not None else []
                cmd += [driver or O.REF_DRIVER, "run", "data", init, str(n), "all%d" % q]
                procs.append(sp.Popen(cmd, cwd=d, stdout=sp.DEVNULL, stderr=sp.DEVNULL, preexec_fn=O._big_stack))
            ok = all(pr.wait(timeout=900) == 0 for pr in procs)
            if ok:
                secs = [float(np.fromfile(os.path.join(d, "all%d" % q, "loop_seconds.f8"))[0]) for q in range(len(cores))]
                res["all_cores_value"] = sum(p.np * 3 * n / x for x in secs)
                res["all_cores_procs"] = len(cores)
                res["all_cores_sample"] = ("%d processes of the same run at once, one per core (%s), rates summed"
                                           % (len(cores), "pinned" if pin is not None else "not pinned"))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return res


def cpu_baseline(name, seconds=12.0):
    """The CPU baseline on a bounded sample of the same workload family (a smaller mesh of the same
    generator, as many steps as fit in about `seconds`): the reference Fortran with the rescale loop
    hoisted (the reference's own arithmetic, bitwise, without its O(np^2) loop) when its build is
    present (oracle/_ref), pinned to one core, with the all-cores figure beside it; then the
    unmodified reference on a smaller mesh and the C restatement for comparison.  Without the
    reference build, the C restatement alone."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as O
    size = {"bumps": 262_144, "dam_break": 561, "paraboloid": 81, "sgn": 1601}[name]
    p = build_problem(name, size)
    port = port_baseline(name, p, seconds)
    if not (O.ref_available() and os.path.exists(O.REF_DRIVER_HOIST)):
        return port
    try:
        ref = reference_baseline(name, p, seconds, driver=O.REF_DRIVER_HOIST)
        # the unmodified reference, whose rate falls with the mesh size (O(np^2) rescale loop), on a
        # mesh a quarter the size, a few seconds
        small = build_problem(name, {"bumps": 65_536, "dam_break": 281, "paraboloid": 81, "sgn": 801}[name])
        plain = reference_baseline(name, small, seconds / 3, all_cores=False)
    except (RuntimeError, OSError, subprocess.SubprocessError) as e:
        print("# reference baseline failed (%s); reporting the C restatement" % str(e)[:200],
              file=sys.stderr, flush=True)
        return port
    ref["unmodified_value"] = plain["value"]
    ref["unmodified_sample"] = plain["sample"]
    ref["port_value"] = port["value"]
    ref["port_sample"] = port["sample"]
    return ref


def port_baseline(name, p, seconds):
    """The C restatement of the reference step (oracle/), 1 thread, on the mesh of problem `p`."""
    import pyoracle as O
    hd, uxd, uyd = p.dirichlet()
    pr = p.params
    visc = {"galerkin": O.GALERKIN, "viscous": O.VISCOUS, "entropy_visc": O.ENTROPY_VISC, "fct": O.FCT}
    orc = O.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=pr["gravity"], htiny=pr["htiny"], cfl=pr["cfl"],
                   max_water_h=pr["max_water_h"], viscosity_type=visc[pr["viscosity"]],
                   viscous_type=pr["viscous_type"], alpha_limit=pr["alpha_limit"], type_test=pr["type_test"],
                   mannings=pr["mannings"], lambda_sgn=pr.get("lambda_sgn", 0.0))
    kind, amp, period, level = p.bc
    kw = dict(bc_kind=kind, amp=amp, period=period, level=level, nodes_h=hd, nodes_ux=uxd, nodes_uy=uyd)
    u = p.u0
    t0 = time.perf_counter()
    u1, _, t, err = orc.run(u, 1, **kw)
    one = time.perf_counter() - t0
    n = max(1, int(seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    u2, _, t, err = orc.run(u1, n, t0=t, **kw)
    el = time.perf_counter() - t0
    rate = p.np * 3 * n / el
    return {"value": rate, "unit": "node-stage updates/s", "cores": 1, "kind": "port",
            "sample": "%s, %d nodes, %d steps (%.1f s), CPU restatement of the reference step, 1 thread"
                      % (p.name, p.np, n, el)}


def hbm_ceiling(gib=2, reps=5):
    """The device's achievable HBM bandwidth (GB/s): streaming read and streaming copy kernels of the
    library (swe_stream_bandwidth), the second ceiling the kernels are quoted against."""
    import ctypes
    from swe_amd.solver import _check, load_library
    r, c = ctypes.c_double(), ctypes.c_double()
    _check(load_library().swe_stream_bandwidth(gib << 30, reps, ctypes.byref(r), ctypes.byref(c)))
    return {"read_GBs": r.value, "copy_GBs": c.value, "bytes": gib << 30}


def copy_calibration(gib=1, reps=3):
    """A plain device-to-device copy of a known size through the HIP runtime, to calibrate the
    FETCH_SIZE / WRITE_SIZE counters (each copy reads and writes `gib` GiB)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(gib << 30)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), n) != 0:
            raise SystemExit("hipMalloc failed")
    if hip.hipMemset(a, 1, n) != 0:
        raise SystemExit("hipMemset failed")
    for _ in range(reps):
        if hip.hipMemcpy(b, a, n, 3) != 0:          # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("copied %d x %d GiB device to device" % (reps, gib))


def _batch_caps(ps, spare_rows=0):
    rows = sum((p.np + 63) // 64 * 64 for p in ps) + spare_rows
    entries = int(sum(p.nnz for p in ps) * 1.4) + 64 * 64
    bnd = sum(sum(x.size for x in p.dirichlet()) for p in ps) + 64
    return rows, entries, bnd


def bench_batch(count, steps, warmup, workers, tune=True, seed=20261202, nmin=5000, nmax=500000, dropin_steps=3,
                dropin_count=None, device_setup=True):
    """Batched form: `count` meshes of the 16.8M-node workload's family, sizes log-uniform in
    [nmin, nmax], dam-break or still-water initial data, all resident in one batch and advanced
    together (one captured graph per step, autotuned schedules); against two plain per-instance
    loops over the same meshes one at a time: the device-resident loop (state kept on the device,
    one captured graph per step) and the drop-in loop (the reference's host time loop calling
    compute_dij and euler per stage, `dropin_steps` steps per mesh, on the first `dropin_count`
    meshes, a random sample of the family, when given).  device_setup: the batch and the
    device-resident loop admit each mesh from its element list (assembly and layout on the GPU, with
    the Hilbert locality order) rather than from the host assembly and layout (reverse Cuthill-McKee)."""
    from swe_amd import problems
    from swe_amd.batch import Batch
    t0 = time.perf_counter()
    ps = problems.build_family(problems.batch_family_specs(count, nmin, nmax, seed=seed), workers)
    t_gen = time.perf_counter() - t0
    nodes = sum(p.np for p in ps)
    rows, entries, bnd = _batch_caps(ps)
    print("# batch: %d meshes, %d nodes (%d..%d), generated in %.1f s" % (count, nodes, min(p.np for p in ps),
          max(p.np for p in ps), t_gen), file=sys.stderr, flush=True)
    total = warmup + steps
    b = Batch(rows, entries, bnd, count)
    t0 = time.perf_counter()
    ids = [b.add_problem(p, max_steps=total, device_setup=device_setup) for p in ps]
    t_admit = time.perf_counter() - t0
    if tune:
        b.tune(2)
    b.step(warmup)
    b.sync()
    before = [b.status(k)[2] for k in ids]
    b.enable_kernel_timing("k_fct_lo")
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    el = time.perf_counter() - t0
    nl, kms = b.kernel_timing("k_fct_lo")
    b.enable_kernel_timing(None)
    # the timed steps again without per-kernel events, as one captured graph per step
    stat = [b.status(k) for k in ids]
    work = sum(p.np * 3 * (st[2] - n0) for p, st, n0 in zip(ps, stat, before))
    failed = sum(1 for st in stat if st[3] != 0)
    b.close()
    b = Batch(rows, entries, bnd, count)
    ids = [b.add_problem(p, max_steps=total, device_setup=device_setup) for p in ps]
    if tune:
        b.tune(2)
    b.step(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    el_graph = time.perf_counter() - t0
    b.close()
    # plain loop: one mesh at a time on a batch with room for the largest
    big = max(ps, key=lambda p: p.nnz)
    r1, e1, b1 = _batch_caps([big])
    one = Batch(r1, e1, max(b1, max(sum(x.size for x in p.dirichlet()) for p in ps) + 64), 1)
    el_plain, work_plain = 0.0, 0
    chunk = 10
    for p in ps:
        k = one.add_problem(p, max_steps=total, device_setup=device_setup)
        one.step(warmup)
        one.sync()
        n0 = one.status(k)[2]
        t0 = time.perf_counter()
        left = steps
        while left > 0:                     # in chunks, so a mesh that stops early stops stepping
            one.step(min(chunk, left))
            left -= chunk
            st = one.status(k)
            if st[0] != 1:
                break
        el_plain += time.perf_counter() - t0
        work_plain += p.np * 3 * (one.status(k)[2] - n0)
        one.remove(k)
    one.close()
    # plain drop-in loop: the reference's time loop on the host calling the per-stage entry points
    # (compute_dij, euler) of each mesh, one mesh at a time, state over PCIe every stage
    from swe_amd.dropin import DropinLoop
    el_drop, work_drop = 0.0, 0
    for p in ps[:dropin_count]:
        with p.solver() as s:
            loop = DropinLoop(s, p)
            loop.run(p.u0, 1)
            t0 = time.perf_counter()
            loop.run(p.u0, dropin_steps)
            el_drop += time.perf_counter() - t0
            work_drop += p.np * 3 * dropin_steps
    kavg = sum(p.nnz for p in ps) / nodes
    bpn = kernel_bytes_per_node("k_fct_lo", kavg, True)
    launch_s = kms / 1e3 / max(nl, 1)
    achieved = bpn * nodes / launch_s / 1e9
    rate_b = work / el_graph
    rate_p = work_plain / el_plain
    rate_d = work_drop / el_drop
    return {"instances": count, "nodes": nodes, "sizes": [nmin, nmax], "steps": steps, "warmup": warmup,
            "node_stage_per_s": rate_b, "ms_per_step": el_graph / steps * 1e3,
            "node_stage_per_s_eager_timed": work / el, "plain_node_stage_per_s": rate_p,
            "plain_ms_total": el_plain * 1e3, "speedup_vs_plain": rate_b / rate_p,
            "dropin_node_stage_per_s": rate_d, "dropin_steps": dropin_steps,
            "dropin_meshes": len(ps[:dropin_count]), "speedup_vs_dropin": rate_b / rate_d,
            "failed_instances": failed,
            "admission_s": t_admit, "setup": "device" if device_setup else "host",
            "roofline": {"bound": "hbm", "kernel": "k_fct_lo", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "launch_ms": launch_s * 1e3,
                         "bytes_per_node": bpn}}


def bench_serve(count, rate, workers, chunk=16, budget_gb=200.0, seed=20261202, median=100000, nmax=2000000,
                device_setup=False):
    """Served form: a Poisson stream of `count` requests at `rate` per second (sizes lognormal with
    the given median, capped at nmax; final times of 1-20 wave periods; 20% dry-bed dam breaks)
    admitted into one batch whose capacity is the memory budget, against the same stream served
    one request at a time.  Both runs must give identical results per request."""
    from swe_amd import problems
    from swe_amd.batch import Batch
    from swe_amd.serve import BatchEngine, Request, Server, run_plain, summarize
    specs = problems.serve_request_specs(count, rate, seed=seed, median=median, nmax=nmax)
    t0 = time.perf_counter()
    ps = problems.build_family(specs, workers)
    print("# serve: %d requests at %.0f/s, %d nodes in all (max %d), generated in %.1f s"
          % (count, rate, sum(p.np for p in ps), max(p.np for p in ps), time.perf_counter() - t0),
          file=sys.stderr, flush=True)

    def reqs():
        return [Request(rid=q, problem=p, tfinal=sp["tfinal"], arrival=sp["arrival"])
                for q, (p, sp) in enumerate(zip(ps, specs))]
    budget = budget_gb * (1 << 30)
    need = Batch.bytes_for(*_batch_caps(ps), count)
    b = Batch.for_budget(min(budget, 1.05 * need + (1 << 30)), max_instances=max(64, count))
    served = Server(BatchEngine(b, device_setup=device_setup), chunk=chunk).run(reqs())
    resident = b.caps.max_nodes
    b.close()
    print("# served: makespan %.1f s" % summarize(served)["makespan_s"], file=sys.stderr, flush=True)
    big = max(ps, key=lambda p: p.nnz)
    r1, e1, b1 = _batch_caps([big])
    one = Batch(r1, e1, max(b1, max(sum(x.size for x in p.dirichlet()) for p in ps) + 64), 1)
    plain = run_plain(BatchEngine(one), reqs(), chunk=chunk)
    one.close()
    same = all(a.steps == c.steps and a.time == c.time and np.array_equal(a.state, c.state, equal_nan=True)
               for a, c in zip(served, plain))
    s, pl = summarize(served), summarize(plain)
    return {"requests": count, "rate_per_s": rate, "budget_gb": budget_gb, "batch_rows": resident,
            "admission": "device set-up from the element list" if device_setup else "host layout",
            "chunk_steps": chunk, "p50_s": s["p50_s"], "p99_s": s["p99_s"], "mean_s": s["mean_s"],
            "node_stage_per_s": s["node_stage_per_s"], "makespan_s": s["makespan_s"],
            "plain_p50_s": pl["p50_s"], "plain_p99_s": pl["p99_s"], "plain_node_stage_per_s": pl["node_stage_per_s"],
            "speedup_vs_plain": s["node_stage_per_s"] / pl["node_stage_per_s"],
            "results_identical_to_plain": bool(same), "errors": s["errors"]}


def replay_fifo(arrivals, durations):
    """Finish times of a one-at-a-time FIFO server given arrival times and each request's solve
    duration: request k starts at max(its arrival, the previous finish)."""
    fin, t = [], 0.0
    for a, d in zip(arrivals, durations):
        t = max(a, t) + d
        fin.append(t)
    return fin


def bench_serve_sweep(count, rates, workers, chunk=16, budget_gb=200.0, seed=20261202, median=100000, nmax=2000000,
                      device_setup=True):
    """The served form over several arrival rates: the same `count` requests (sizes lognormal with
    the given median, capped at nmax, final times of 1-20 wave periods, 20% dry-bed dam breaks; the
    seed fixes them, the rate only scales the arrival gaps) served by continuous batching at each
    rate against a 200 GB budget.  The plain loop serves the stream one request at a time once; its
    per-request solve times (admission, steps, read-back) do not depend on arrivals, so its
    finish times at every rate follow from them by the FIFO recurrence (replay_fifo).  Every
    served result must equal the plain loop's."""
    from swe_amd import problems
    from swe_amd.batch import Batch
    from swe_amd.serve import BatchEngine, Request, Server, run_plain, summarize
    specs = {r: problems.serve_request_specs(count, r, seed=seed, median=median, nmax=nmax) for r in rates}
    base = specs[rates[-1]]
    t0 = time.perf_counter()
    ps = problems.build_family(base, workers)
    print("# serve: %d requests, %d nodes in all (max %d), generated in %.1f s"
          % (count, sum(p.np for p in ps), max(p.np for p in ps), time.perf_counter() - t0), file=sys.stderr,
          flush=True)

    def reqs(sp):
        return [Request(rid=q, problem=p, tfinal=x["tfinal"], arrival=x["arrival"]) for q, (p, x) in
                enumerate(zip(ps, sp))]
    # the plain loop, at the fastest rate (no idle time between requests)
    big = max(ps, key=lambda p: p.nnz)
    r1, e1, b1 = _batch_caps([big])
    one = Batch(r1, e1, max(b1, max(sum(x.size for x in p.dirichlet()) for p in ps) + 64), 1)
    plain = run_plain(BatchEngine(one, device_setup=device_setup), reqs(base), chunk=chunk)
    one.close()
    plain_by_id = {r.rid: r for r in plain}
    dur = [plain_by_id[q].finished - max(plain_by_id[q].admitted, 0.0) for q in range(count)]
    work = float(sum(r.node_stages for r in plain))
    budget = budget_gb * (1 << 30)
    need = Batch.bytes_for(*_batch_caps(ps), count)
    out = {"requests": count, "budget_gb": budget_gb, "chunk_steps": chunk, "nodes": int(sum(p.np for p in ps)),
           "max_nodes": int(max(p.np for p in ps)),
           "dry_bed_fraction": sum(1 for x in base if x["kind"] == "drydam") / float(count),
           "setup": "device" if device_setup else "host",
           "plain_solve_s": float(sum(dur)), "rates": []}
    for rate in rates:
        sp = specs[rate]
        b = Batch.for_budget(min(budget, 1.05 * need + (1 << 30)), max_instances=max(64, count))
        served = Server(BatchEngine(b, device_setup=device_setup), chunk=chunk).run(reqs(sp))
        b.close()
        same = all(a.steps == plain_by_id[a.rid].steps and a.time == plain_by_id[a.rid].time and
                   np.array_equal(a.state, plain_by_id[a.rid].state, equal_nan=True) for a in served)
        arr = [x["arrival"] for x in sp]
        fin = replay_fifo(arr, dur)
        lat = np.array([f - a for f, a in zip(fin, arr)])
        span = max(fin) - min(arr)
        sv = summarize(served)
        row = {"rate_per_s": rate, "p50_s": sv["p50_s"], "p99_s": sv["p99_s"], "mean_s": sv["mean_s"],
               "node_stage_per_s": sv["node_stage_per_s"], "makespan_s": sv["makespan_s"],
               "plain_p50_s": float(np.percentile(lat, 50)), "plain_p99_s": float(np.percentile(lat, 99)),
               "plain_node_stage_per_s": work / span, "speedup_vs_plain": sv["node_stage_per_s"] / (work / span),
               "results_identical_to_plain": bool(same), "errors": sv["errors"]}
        out["rates"].append(row)
        print("# served at %.0f/s: p50 %.2f s p99 %.2f s, %.3g node-stage/s (plain %.3g), identical %s"
              % (rate, row["p50_s"], row["p99_s"], row["node_stage_per_s"], row["plain_node_stage_per_s"], same),
              file=sys.stderr, flush=True)
    out["speedup_vs_plain"] = min(r["speedup_vs_plain"] for r in out["rates"])
    out["p99_s"] = max(r["p99_s"] for r in out["rates"])
    return out


def bench_setup(n_side=4097, reps=2):
    """P1 set-up (construct_matrices) of the 16.8M-node mesh: host assembly (all host threads)
    against the device assembly (including its host<->device copies); both must give the same
    bits."""
    from swe_amd import mesh as M
    from swe_amd.solver import assemble
    t0 = time.perf_counter()
    m = M.renumber_first_touch(M.jittered_delaunay_mesh(n_side, 1000.0, jitter=0.25, seed=20261202))
    print("# setup: mesh of %d nodes, %d elements generated in %.1f s" % (m.np, m.jj.shape[0], time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    assemble(np.array([[0, 1, 2]]), np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]), device=True)   # warm-up
    t_host, t_dev = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        host = assemble(m.jj, m.rr)
        t_host.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        dev = assemble(m.jj, m.rr, device=True)
        t_dev.append(time.perf_counter() - t0)
        same = all(np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)) for a, b in zip(host, dev))
        del host, dev
        print("# setup: host %.2f s, device %.2f s, identical %s" % (t_host[-1], t_dev[-1], same), file=sys.stderr,
              flush=True)
    # the whole set-up of a solver: host assembly + host layout (swe_create) against device assembly +
    # layout built on the GPU (swe_create_from_assembly); then two steps of each must agree bitwise
    from swe_amd import problems as PR
    from swe_amd.solver import BC_INFLOW, Solver
    bath = PR.bump_field(m.rr, 1000.0, 64, 20261202)
    prm = dict(gravity=9.81, htiny=1.0e-4, cfl=0.25, max_water_h=1.0, viscosity="fct", viscous_type=2,
               alpha_limit=True, type_test=4, mannings=0.0)
    lists = dict(h_dir=M.dirichlet_nodes(m, (4,)), ux_dir=M.dirichlet_nodes(m, (2,)), uy_dir=M.dirichlet_nodes(m, (1, 3)))
    u0 = np.zeros((m.np, 3))
    u0[:, 0] = np.maximum(1.0 - bath, 0.0)
    res = []
    for dev_path in (False, True):
        t0 = time.perf_counter()
        if dev_path:
            s = Solver.from_elements(m.jj, m.rr, bath, **lists, **prm)
        else:
            ia, ja, _diag, _mass, lumped, cij = assemble(m.jj, m.rr)
            s = Solver(ia, ja, cij, lumped, bath, **lists, **prm)
            del ia, ja, _diag, _mass, lumped, cij
        dt_create = time.perf_counter() - t0
        s.set_bc(BC_INFLOW, 0.05, 20.0, 1.0)
        s.set_state(u0)
        _t, dts = s.run(2, time=0.0)
        res.append((dt_create, s.get_state(), dts))
        s.close()
        del s
        print("# setup: solver created (%s) in %.2f s" % ("device" if dev_path else "host", dt_create),
              file=sys.stderr, flush=True)
    same_run = bool(np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2]))
    return {"nodes": int(m.np), "elements": int(m.jj.shape[0]), "host_s": min(t_host), "device_s": min(t_dev),
            "host_threads": os.cpu_count(), "identical": bool(same),
            "solver_host_s": res[0][0], "solver_device_s": res[1][0], "solver_runs_identical": same_run}


def bench_errors(n=4097, reps=5):
    """compute_errors (the analytic-error check of the paraboloid, test 5) on an n x n grid of
    [0,4]^2 after 2 steps: the device evaluator (kernel plus the 96-byte read-back) against the
    oracle's sequential sums on one host core; the two agree within the sequential sums' own
    rounding bound (the device sums are the exactly rounded ones)."""
    from swe_amd import problems as PR
    from swe_amd import mesh as M
    from swe_amd.solver import Solver
    t0 = time.perf_counter()
    m = M.structured_mesh(n, n, 0.0, 4.0, 0.0, 4.0)
    g, h0, a, hL = 9.81, 0.1, 1.0, 4.0
    htiny = 1.0e-4 * hL
    bath = -h0 * (1.0 - ((m.rr[:, 0] - hL / 2.0) ** 2 + (m.rr[:, 1] - hL / 2.0) ** 2) / a ** 2)
    u0 = PR.paraboloid_state(m.rr, 0.0, htiny)
    print("# errors: grid of %d nodes, %d elements in %.1f s" % (m.np, m.jj.shape[0], time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    s = Solver.from_elements(m.jj, m.rr, bath, gravity=g, htiny=htiny, cfl=0.25, max_water_h=hL,
                             viscosity="viscous", type_test=5)
    s.set_state(u0)
    s.run(2, time=0.0)
    ev = s.error_evaluator(m.jj, m.rr)
    ev.sums()
    ts = []
    for _ in range(reps):
        t1 = time.perf_counter()
        dev = ev.sums()
        ts.append(time.perf_counter() - t1)
    u, t = s.get_state(), s.time
    rel = ev.relative()
    ev.close()
    s.close()
    res = {"nodes": int(m.np), "elements": int(m.jj.shape[0]), "device_s": min(ts), "l1_h": rel["l1_h"],
           "l2_h": rel["l2_h"]}
    # the CPU baseline leg of this workload: the oracle's sequential sums on one host core, timed,
    # and used as the checker of the device sums
    try:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle
        t1 = time.perf_counter()
        ref = pyoracle.compute_errors(m.jj, m.rr, u, 5, t, g, htiny)
        cpu = {"value": time.perf_counter() - t1, "unit": "s", "cores": 1, "kind": "port",
               "sample": "the same sums on the same state (C restatement of compute_errors, 1 thread)"}
        cpu["max_rel_diff"] = float(np.max(np.abs(dev - ref) / np.maximum(np.abs(ref), 1e-300)))
        # the summands are >= 0, so the sequential sums are within (n-1)u/(1-(n-1)u) of the exact
        # ones, relative (half that after the square root of the L2 sums)
        nq = 3 * m.jj.shape[0]
        cpu["sequential_rounding_bound"] = (nq - 1) * 2.0 ** -53 / (1 - (nq - 1) * 2.0 ** -53)
        cpu["within_bound"] = bool(cpu["max_rel_diff"] <= cpu["sequential_rounding_bound"])
        res["cpu_baseline"] = cpu
    except (ImportError, OSError) as e:
        res["cpu_baseline"] = "not available: %s" % e
    print("# errors: device %.4f s; %s" % (res["device_s"], {k: v for k, v in res.items() if k != "device_s"}),
          file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="bumps",
                    choices=["bumps", "bumps_lattice", "dam_break", "paraboloid", "sgn", "batch", "serve",
                             "setup", "errors"])
    ap.add_argument("--instances", type=int, default=256, help="batch: meshes in the batch")
    ap.add_argument("--requests", type=int, default=64, help="serve: requests in the stream")
    ap.add_argument("--rates", default="20", help="serve: arrival rates (per s), comma separated")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: device memory budget")
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="processes that generate the meshes of the batch and serve workloads")
    ap.add_argument("--device-setup", action="store_true",
                    help="serve: admit requests from their element lists (set-up and layout on the GPU)")
    ap.add_argument("--no-extra", action="store_true",
                    help="default run: leave out the reduced batch and serve measurements")
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch schedules")
    ap.add_argument("--step-form", type=int, default=None,
                    help="force the step form after tuning (0 kernel launches / graph, 1 persistent)")
    ap.add_argument("--lo-kind", type=int, default=None,
                    help="force the schedule kind of k_fct_lo after tuning (0 lambda inline, 1 lambda in LDS)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy used to calibrate the HBM byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding of a single mesh is a design note "
              "only, so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from swe_amd.solver import load_library
    load_library()
    if args.workload == "batch":
        res = {"metric": "node-stage updates/s (FP64, SSP-RK3, FCT + alpha), batched", "unit": "node-stage updates/s",
               "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic"}
        r = bench_batch(args.instances, args.steps, args.warmup, args.workers, tune=not args.no_tune)
        res.update({"value": r["node_stage_per_s"], "steps": args.steps, "warmup": args.warmup,
                    "ms_per_step": r["ms_per_step"], "roofline": r.pop("roofline"),
                    "config": {"workload": "batch", **r}})
        print(json.dumps(res))
        return
    if args.workload == "errors":
        r = bench_errors()
        print(json.dumps({"metric": "compute_errors time (analytic L1/L2 check), device", "value": r["device_s"],
                          "unit": "s", "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": "errors", **r}}), flush=True)
        return
    if args.workload == "setup":
        r = bench_setup()
        print(json.dumps({"metric": "P1 set-up time (construct_matrices), device", "value": r["device_s"], "unit": "s",
                          "n_gpus": 1, "higher_is_better": False, "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic", "config": {"workload": "setup", **r}}), flush=True)
        return
    if args.workload == "serve":
        for rate in [float(x) for x in args.rates.split(",")]:
            r = bench_serve(args.requests, rate, args.workers, budget_gb=args.budget_gb,
                            device_setup=args.device_setup)
            print(json.dumps({"metric": "time-to-solution p99 (served, FP64 FCT + alpha)", "value": r["p99_s"],
                              "unit": "s", "n_gpus": 1, "higher_is_better": False, "scaling": "weak",
                              "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                              "config": {"workload": "serve", **r}}), flush=True)
        return
    t0 = time.perf_counter()
    # the unstructured workload is assembled and laid out on the GPU (swe_create_from_assembly)
    device_setup = args.workload == "bumps"
    p = build_problem(args.workload, args.size, host_assembly=not device_setup)
    t_gen = time.perf_counter() - t0
    s = p.solver(device_setup=device_setup)
    t_setup = time.perf_counter() - t0
    nnz = s.nnz
    k = nnz / p.np
    fct = p.params["viscosity"] == "fct"
    nc = 5 if p.params.get("type_test") == 13 else 3
    alpha = p.params["alpha_limit"]
    dominant = "k_fct_lo" if fct else "k_stage_lo"
    print("# workload %s: %d nodes, %d entries (k=%.3f), mesh and data %.1f s, solver %.1f s"
          % (p.name, p.np, nnz, k, t_gen, t_setup - t_gen), file=sys.stderr, flush=True)
    s.set_state(p.u0)
    s.time = 0.0
    if not args.no_tune:
        s.tune(3)           # launch schedules (workgroup size, XCD order, kind) timed on this mesh; state restored
    if args.lo_kind is not None:
        s.schedule = {n: (b, r, args.lo_kind if n == "k_fct_lo" else kd) for n, (b, r, kd) in s.schedule.items()}
    if args.step_form is not None:
        s.step_form = args.step_form
    s.run(args.warmup)
    s.sync()
    s.enable_kernel_timing(dominant)
    t0 = time.perf_counter()
    s.run_async(args.steps)
    s.sync()
    el = time.perf_counter() - t0
    nlaunch, kms = s.kernel_timing(dominant)
    s.enable_kernel_timing(None)
    hmin = float(s.get_state()[:, 0].min())
    schedule = {k: list(v) for k, v in s.schedule.items()}
    # every kernel of the step against the same roofline (3 more steps per kernel, after the timed ones)
    per_kernel = {}
    names = (["k_compute_dt", "k_alpha", "k_fct_lo", "k_fct_rminus", "k_fct_apply"] if fct else
             ["k_compute_dt", "k_alpha", "k_stage_lo"] if alpha else ["k_compute_dt", "k_stage_lo"])
    for name in names:
        if name == dominant:
            n_l, t_ms = nlaunch, kms
        else:
            s.enable_kernel_timing(name)
            s.run_async(3)
            s.sync()
            n_l, t_ms = s.kernel_timing(name)
            s.enable_kernel_timing(None)
        if n_l:
            ms_l = t_ms / n_l
            gbs = kernel_bytes_per_node(name, k, alpha, nc) * p.np / (ms_l / 1e3) / 1e9
            per_kernel[name] = {"launch_ms": ms_l, "per_step": n_l / (args.steps if name == dominant else 3),
                                "alg_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    rate = p.np * 3 * args.steps / el
    bpn = kernel_bytes_per_node(dominant, k, alpha, nc)
    avg_s = kms / 1e3 / max(nlaunch, 1)
    achieved = bpn * p.np / avg_s / 1e9
    ceiling = hbm_ceiling()
    achievable = max(ceiling["read_GBs"], ceiling["copy_GBs"])
    for v in per_kernel.values():
        v["frac_of_achievable"] = v["alg_GBs"] / achievable
    res = {
        "metric": "node-stage updates/s (FP64, SSP-RK3, FCT + alpha)" if fct else "node-stage updates/s (FP64, SSP-RK3)",
        "value": rate,
        "unit": "node-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": p.name, "mesh": MESH_NAMES.get(args.workload, ""), "nodes": p.np, "entries": nnz,
                   "viscosity": p.params["viscosity"], "alpha_limit": bool(alpha), "steps": args.steps,
                   "hmin_end": hmin, "schedule": schedule, "step_form": s.step_form,
                   "setup": "device" if device_setup else "host"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant, p.name, p.np),
                     "valu_busy_and_lanes": measured_valu(dominant, p.name, p.np),
                     "launch_ms": avg_s * 1e3, "bytes_per_node": bpn, "achievable_GBs": achievable,
                     "frac_of_achievable": achieved / achievable},
        "hbm_ceiling": ceiling,
        "step_effective_GBs": rate * step_bytes_per_node_stage(k, fct, alpha, nc) / 1e9,
        "step_frac": rate * step_bytes_per_node_stage(k, fct, alpha, nc) / 1e9 / HBM_PEAK_GBS,
        "step_frac_of_achievable": rate * step_bytes_per_node_stage(k, fct, alpha, nc) / 1e9 / achievable,
        "kernels": per_kernel,
    }
    s.close()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
    if args.workload == "bumps" and args.size is None and not args.no_extra:
        # the batched form at its configuration (256 meshes of 5k-500k nodes) and the served form's
        # arrival-rate sweep, each against the plain per-instance loop
        res["batched"] = bench_batch(256, 30, 3, args.workers, dropin_count=32)
        res["served"] = bench_serve_sweep(32, (5.0, 50.0, 200.0), args.workers)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
