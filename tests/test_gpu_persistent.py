"""The persistent step form (SWE_STEP_PERSISTENT: the fct scheme's passes in one cooperative kernel
with grid-wide barriers) against the reference fixtures and against the launch sequence.

Bar: the same bits as the reference's runs (Manning runs within 1e-11, as the launch form), and the
same bits as the launch form on every instance of a batch, including instances that stop in the
middle of a launch of many steps."""
import numpy as np
import pytest

from conftest import golden, rel_diff, same_bits

pytestmark = pytest.mark.gpu

TOL = 1e-11
FCT_RUNS = {
    "dambreak_run.npz": (True, 10, 0.02, 1),
    "dambreak9_run.npz": (True, 9, 0.02, 1),
    "bumps_run.npz": (True, 4, 0.0, 2),
    "cav025_run.npz": (True, 4, 0.0, 0),
}


def fct_solver(z, alpha, type_test, mannings, bc):
    from swe_amd.solver import Solver
    g, htiny, hmax, cfl = z["scalars"]
    s = Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], gravity=g, htiny=htiny, cfl=cfl,
               max_water_h=hmax, viscosity="fct", alpha_limit=alpha, type_test=type_test, mannings=mannings,
               h_dir=z["h_dir"], ux_dir=z["ux_dir"], uy_dir=z["uy_dir"])
    if bc == 2:
        amp, period, level = z["bc"]
        s.set_bc(2, amp, period, level)
    else:
        s.set_bc(bc)
    return s


@pytest.mark.parametrize("name", sorted(FCT_RUNS))
def test_persistent_runs_against_reference(name):
    from swe_amd.solver import STEP_PERSISTENT
    z = golden(name)
    alpha, tt, mannings, bc = FCT_RUNS[name]
    with fct_solver(z, alpha, tt, mannings, bc) as s:
        s.step_form = STEP_PERSISTENT
        assert s.step_form == STEP_PERSISTENT
        s.set_state(z["u0"])
        t, done = float(z["time0"]), 0
        for snap, step in zip(z["snaps"], z["snap_steps"]):
            print(name, "persistent to step", step, flush=True)
            t, d = s.run(int(step) - done, time=t)
            u = s.get_state()
            if mannings > 0:
                l1, linf = rel_diff(u, snap)
                assert l1 <= TOL and linf <= TOL, (step, l1, linf)
            else:
                assert same_bits(d, z["dts"][done:int(step)])
                assert same_bits(u, snap), (name, step)
            done = int(step)
        t, d = s.run(z["dts"].size - done, time=t)
        u = s.get_state()
        if mannings == 0:
            assert same_bits(u, z["final"]) and t == float(z["time"])
        else:
            l1, linf = rel_diff(u, z["final"])
            assert l1 <= TOL and linf <= TOL


def test_persistent_batch_equals_launch_form():
    """Four meshes with different step limits in one batch, stepped 40 at a time in the persistent
    form (so instances stop inside a launch): every instance equals the launch form's batch."""
    from swe_amd import problems
    from swe_amd.batch import Batch
    from swe_amd.solver import STEP_LAUNCHES, STEP_PERSISTENT
    ps = [problems.dam_break(101, 21, seed=7), problems.bumps(n_side=65, L=1000.0, seed=11),
          problems.dam_break(61, 13, seed=3), problems.bumps(n_side=33, L=500.0, seed=5, period=7.0)]
    limits = [40, 17, 33, 25]
    res = []
    for form in (STEP_LAUNCHES, STEP_PERSISTENT):
        nodes = sum((p.np + 63) // 64 * 64 for p in ps) * 2
        entries = int(sum(p.nnz for p in ps) * 1.4) * 2 + 4096
        bnd = sum(sum(x.size for x in p.dirichlet()) for p in ps) * 2 + 64
        with Batch(nodes, entries, bnd, 16) as b:
            b.step_form = form
            ids = [b.add_problem(p, max_steps=n) for p, n in zip(ps, limits)]
            print("batch, step form", form, flush=True)
            b.step(40)
            b.step(2)
            res.append([(b.status(k), b.get_state(k)) for k in ids])
    for (sa, ua), (sb, ub), n in zip(res[0], res[1], limits):
        assert sa == sb and sa[2] == n and sa[3] == 0
        assert same_bits(ua, ub)


def test_tuner_picks_a_form_and_results_stay():
    """swe_tune times both step forms; whichever it keeps, the run gives the same bits."""
    from swe_amd import problems
    p = problems.bumps(n_side=65, L=1000.0, seed=11)
    with p.solver() as s:
        s.set_state(p.u0)
        ref_t, _ = s.run(20, time=0.0)
        ref = s.get_state()
    with p.solver() as s:
        s.set_state(p.u0)
        sched = s.tune(steps=2)
        print("tuned", sched, "step form", s.step_form, flush=True)
        assert s.time == 0.0 and same_bits(s.get_state(), p.u0)
        t, _ = s.run(20, time=0.0)
        assert t == ref_t and same_bits(s.get_state(), ref)


def test_step_form_refusals():
    from swe_amd import problems
    from swe_amd.solver import SWE_EINVAL, STEP_LAUNCHES, SweError
    p = problems.bumps(n_side=33, L=500.0, seed=5)
    with p.solver() as s:
        with pytest.raises(SweError) as e:
            s.step_form = 7
        assert e.value.status == SWE_EINVAL
        assert s.step_form == STEP_LAUNCHES


def test_barrier_that_gives_up_fails_the_step(monkeypatch):
    """A barrier whose wait bound is forced to zero from the host (SWE_BARRIER_SPINS=0): the workgroups
    leave the persistent kernel at once, and the run reports SWE_EHIP instead of a state; a new solver
    with the normal bound then runs bitwise as before."""
    from swe_amd import problems
    from swe_amd.solver import STEP_PERSISTENT, SWE_EHIP, SweError
    p = problems.bumps(n_side=65, L=1000.0, seed=11)
    with p.solver() as s:
        s.set_state(p.u0)
        t_ref, _ = s.run(5, time=0.0)
        ref = s.get_state()
    monkeypatch.setenv("SWE_BARRIER_SPINS", "0")
    with p.solver() as s:
        s.step_form = STEP_PERSISTENT
        s.set_state(p.u0)
        with pytest.raises(SweError) as e:
            s.run(5, time=0.0)
        assert e.value.status == SWE_EHIP
    monkeypatch.delenv("SWE_BARRIER_SPINS")
    with p.solver() as s:
        s.step_form = STEP_PERSISTENT
        s.set_state(p.u0)
        t, _ = s.run(5, time=0.0)
        assert t == t_ref and same_bits(s.get_state(), ref)
