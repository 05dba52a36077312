"""compute_errors on the device (swe_compute_errors) against the oracle, which is bitwise the
reference's own routines on these states (test_errors_oracle.py).  The device sums are double-double
over a fixed partition, i.e. the exactly rounded sums: they must equal math.fsum of the oracle's
summands to 2 ulp, and the reference's sequential sums within that sum's rounding bound (far below
1e-11 relative on these meshes).  The same holds for swe_ns_l1_ratio, the time loop's unit-10
diagnostic of tests 1, 2 and 10 (main.f90:94-99), against the reference's ns_l1 (tests/golden/flow.npz)."""
import math

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

CASES = {"par41_modes": ("viscous", False, 5), "cav025_run": ("fct", True, 4),
         "cdisk_coriolis": ("viscous", True, 12), "par_friction": ("viscous", True, 11),
         "sgn_run": ("viscous", True, 13)}
# sol_anal through a device libm function (cosh for the SGN wave of test 13): a summand may differ
# from the host's in its last bits, so the sums are compared to the reference's within 1e-12
DEVICE_LIBM = {13}


def _solver(z, viscosity, alpha, test):
    from swe_amd.solver import Solver
    g, htiny, hmax, cfl = z["scalars"]
    return Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], gravity=g, htiny=htiny, cfl=cfl,
                  max_water_h=hmax, viscosity=viscosity, alpha_limit=alpha, type_test=test,
                  h_dir=z["h_dir"], ux_dir=z["ux_dir"], uy_dir=z["uy_dir"])


def _check(dev, oracle_out, terms, ref=None):
    n = terms.shape[-1]
    gam = (n - 1) * 2.0 ** -53 / (1 - (n - 1) * 2.0 ** -53)
    for k in range(3):
        for q in range(4):
            exact = math.fsum(terms[k, q])
            exact = math.sqrt(exact) if q >= 2 else exact
            assert abs(dev[k, q] - exact) <= 4 * 2.0 ** -53 * abs(exact), (k, q, dev[k, q], exact)
            bound = gam * np.abs(terms[k, q]).sum()
            if q >= 2:
                bound = bound / (2 * max(exact, 1e-300)) + 4 * 2.0 ** -53 * exact
            assert abs(dev[k, q] - oracle_out[k, q]) <= bound + 4 * 2.0 ** -53 * abs(exact)
            if ref is not None:
                assert abs(dev[k, q] - ref[k, q]) <= 1e-11 * max(abs(ref[k, q]), 1e-300)


@pytest.mark.parametrize("key", sorted(golden("errors.npz")))
def test_device_errors_match_the_reference(oracle, key):
    name, st, t = key.split("|")
    z = golden(name + ".npz")
    u = z["states"][0] if st == "states0" else z["states"][1] if st == "states1" else z["final"]
    t = float(z["time"]) if t == "time" else float(t)
    print(key, flush=True)
    with _solver(z, *CASES[name]) as s:
        s.set_state(u)
        ev = s.error_evaluator(z["jj"], z["rr"])
        dev = ev.sums(t)
        again = ev.sums(t)
        ev.close()
    g, htiny = z["scalars"][0], z["scalars"][1]
    out, terms = oracle.compute_errors(z["jj"], z["rr"], u, int(z["type_test"]), t, g, htiny, terms=True)
    assert np.array_equal(dev, again)
    ref = golden("errors.npz")[key]
    if int(z["type_test"]) in DEVICE_LIBM:
        scale = np.maximum(np.abs(ref), 1e-300)
        assert np.all(np.abs(dev - ref) <= 1e-12 * scale) and np.all(np.abs(dev - out) <= 1e-12 * scale)
    else:
        _check(dev, out, terms, ref)


def test_device_errors_after_a_run_at_scale(oracle):
    """The paraboloid (test 5) on a 301x301 grid after 50 steps, at the solver's own time; the same
    bits from a solver laid out on the device (another internal numbering)."""
    from swe_amd import problems as PR
    p = PR.paraboloid(n=301)
    print("paraboloid", p.np, flush=True)
    res = []
    for dev_setup in (False, True):
        with p.solver(device_setup=dev_setup) as s:
            s.set_state(p.u0)
            s.run(50, time=0.0)
            u = s.get_state()
            t = s.time
            ev = s.error_evaluator(p.mesh.jj, p.mesh.rr)
            res.append((ev.sums(), ev.relative(), u, t))
            ev.close()
    assert np.array_equal(res[0][0], res[1][0])
    dev, rel, u, t = res[0]
    prm = p.params
    out, terms = oracle.compute_errors(p.mesh.jj, p.mesh.rr, u, 5, t, prm["gravity"], prm["htiny"], terms=True)
    _check(dev, out, terms)
    assert abs(rel["l1_h"] - out[0, 0] / out[0, 1]) <= 1e-11 * rel["l1_h"]


def test_device_errors_reject_tests_without_sol_anal_and_bad_lists():
    from swe_amd.solver import SweError
    z = golden("dambreak_run.npz")            # test 10: no analytic solution
    with _solver(z, "fct", True, 10) as s:
        s.set_state(z["u0"])
        ev = s.error_evaluator(z["jj"], z["rr"])
        with pytest.raises(SweError):
            ev.sums(1.0)
        ev.close()
        bad = np.array(z["jj"], copy=True)
        bad[0, 0] = z["rr"].shape[0]          # node out of range
        with pytest.raises(SweError):
            s.error_evaluator(bad, z["rr"])


def test_batch_instance_errors_equal_solver_alone():
    """swe_batch_errors_create: the sums of a batch instance after a run in a mixed batch, and its
    unit-10 diagnostic, are those of the same mesh run alone (same state bits, same element
    partition); removed instances are refused."""
    from swe_amd import problems as PR
    from swe_amd.batch import Batch
    from swe_amd.solver import SweError
    pa, pb = PR.paraboloid(n=41), PR.paraboloid(n=61)
    pd = PR.dam_break(101, 21, seed=7, viscosity="viscous", alpha=False)    # the pool runs one scheme
    n = 30
    alone, alone_flow = [], []
    for p in (pa, pb):
        with p.solver() as s:
            s.set_state(p.u0)
            s.run(n, time=0.0)
            ev = s.error_evaluator(p.mesh.jj, p.mesh.rr)
            alone.append(ev.sums())
            alone_flow.append(ev.ns_l1_ratio())
            ev.close()
    with Batch(200000, 1500000, 20000, 8) as b:
        ka = b.add_problem(pa, max_steps=n)
        kd = b.add_problem(pd, max_steps=n)
        kb = b.add_problem(pb, max_steps=n)
        b.step(n)
        b.sync()
        eva = b.error_evaluator(ka, pa.mesh.jj, pa.mesh.rr)
        evb = b.error_evaluator(kb, pb.mesh.jj, pb.mesh.rr)
        assert np.array_equal(eva.sums(), alone[0]) and np.array_equal(evb.sums(), alone[1])
        assert np.array_equal(eva.ns_l1_ratio(), alone_flow[0]) and np.array_equal(evb.ns_l1_ratio(), alone_flow[1])
        b.remove(ka)
        with pytest.raises(SweError):
            eva.sums(1.0)
        eva.close()
        evb.close()
        evd = b.error_evaluator(kd, pd.mesh.jj, pd.mesh.rr)     # test 10: no sol_anal
        with pytest.raises(SweError):
            evd.sums(1.0)
        evd.close()


@pytest.mark.parametrize("test", [3, 6, 7])
def test_device_bump_flow_errors(oracle, test):
    """Tests 3, 6, 7 on the reference's strip fixture: at t = 0 the summands are bitwise the oracle's
    (2 ulp from math.fsum); at t = 1 the device's acos/cos/pow may move a summand by an ulp, so the sums
    are held to 1e-12 relative of the reference's."""
    from swe_amd.solver import Solver, assemble
    z = golden("errors_bump.npz")
    g, htiny = z["scalars_%d" % test]
    ia, ja, _d, _m, lumped, cij = assemble(z["jj"], z["rr"])
    s = Solver(ia, ja, cij, lumped, np.zeros(z["rr"].shape[0]), gravity=g, htiny=htiny, cfl=0.25,
               max_water_h=2.0, viscosity="viscous", type_test=test)
    s.set_state(z["u0_%d" % test])
    ev = s.error_evaluator(z["jj"], z["rr"])
    for t in (0, 1):
        dev = ev.sums(float(t))
        ref = z["errors_%d_t%d" % (test, t)]
        if t == 0:
            out, terms = oracle.compute_errors(z["jj"], z["rr"], z["u0_%d" % test], test, 0.0, g, htiny, terms=True)
            _check(dev, out, terms, ref)
        else:
            assert np.all(np.abs(dev - ref) <= 1e-12 * np.maximum(np.abs(ref), 1e-300)), (dev, ref)
    ev.close()
    s.close()


FLOW_CASES = {"dambreak_run": ("fct", True, 10), "sgn_run": ("viscous", True, 13)}


def _flow_check(dev, out, terms, ref=None):
    """dev: the device's [q1, q2, q3, ratio]; exactly rounded sums (math.fsum of the oracle's summands)
    to 4 ulp, the sequential sums within their rounding bound, the ratio to 1e-12."""
    n = terms.shape[-1]
    gam = (n - 1) * 2.0 ** -53 / (1 - (n - 1) * 2.0 ** -53)
    for f in range(3):
        exact = math.fsum(terms[f])
        assert abs(dev[f] - exact) <= 4 * 2.0 ** -53 * abs(exact), (f, dev[f], exact)
        assert abs(dev[f] - out[f]) <= gam * terms[f].sum() + 4 * 2.0 ** -53 * abs(exact)
        if ref is not None:
            assert abs(dev[f] - ref[f]) <= 1e-12 * max(abs(ref[f]), 1e-300)
    assert abs(dev[3] - out[3]) <= 1e-12 * abs(out[3])
    if ref is not None:
        assert abs(dev[3] - ref[3]) <= 1e-12 * abs(ref[3])


@pytest.mark.parametrize("key", sorted(golden("flow.npz")))
def test_device_flow_ratio_matches_the_reference(oracle, key):
    """swe_ns_l1_ratio, the time loop's unit-10 diagnostic (main.f90:94-99), on the dam-break states
    (test 10) and the SGN canal's final state, against the reference's own ns_l1 (tests/golden/flow.npz)."""
    name, st = key.split("|")
    z = golden(name + ".npz")
    u = z["snaps"][list(z["snap_steps"]).index(int(st[4:]))] if st.startswith("snap") else z[st]
    print(key, flush=True)
    with _solver(z, *FLOW_CASES[name]) as s:
        s.set_state(u)
        ev = s.error_evaluator(z["jj"], z["rr"])
        dev = ev.ns_l1_ratio()
        again = ev.ns_l1_ratio()
        ev.close()
    assert np.array_equal(dev, again)
    out, terms = oracle.ns_l1_ratio(z["jj"], z["rr"], u, float(z["scalars"][0]), terms=True)
    _flow_check(dev, out, terms, golden("flow.npz")[key])


def test_device_flow_ratio_every_step_of_a_dam_break(oracle):
    """The loop of main.f90:71-99 for test 10 with the state on the device: one swe_run step, then the
    diagnostic, 40 times on a 301x61 dam break (FCT + alpha + Manning, walls), for a solver laid out on
    the host and one assembled on the device (other internal numbering): the same bits from both, and
    each value against the oracle's ns_l1 of the state the step left."""
    from swe_amd import problems as PR
    p = PR.dam_break(301, 61, seed=5)
    g = p.params["gravity"]
    series = []
    for dev_setup in (False, True):
        with p.solver(device_setup=dev_setup) as s:
            s.set_state(p.u0)
            ev = s.error_evaluator(p.mesh.jj, p.mesh.rr)
            vals, t = [], 0.0
            for step in range(40):
                t, _ = s.run(1, time=t)
                vals.append(ev.ns_l1_ratio())
                if not dev_setup and step % 10 == 9:
                    out, terms = oracle.ns_l1_ratio(p.mesh.jj, p.mesh.rr, s.get_state(), g, terms=True)
                    _flow_check(vals[-1], out, terms)
            ev.close()
        series.append(np.array(vals))
    print("dam break 301x61: ratio after 1 and 40 steps %.6g %.6g" % (series[0][0, 3], series[0][-1, 3]), flush=True)
    assert np.array_equal(series[0], series[1])
    assert np.all(np.diff(series[0][:, 3]) > 0)     # the dam-break front accelerates from rest


def test_device_loop_unit10_series_equals_the_reference_program():
    """Test 1 (FCT + alpha) device-resident from the reference program's own set-up: swe_run one step at
    a time and swe_ns_l1_ratio after each, as a device-resident main.f90 would write unit 10.  Every
    step's time equals the reference program's bit for bit, every ratio its fort.10 value within 1e-11
    (tests/golden/unit10_test1.npz, 268 steps)."""
    from swe_amd.solver import Solver
    z = golden("unit10_test1.npz")
    g, htiny, hmax, cfl = z["scalars"]
    with Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], gravity=g, htiny=htiny, cfl=cfl,
                max_water_h=hmax, viscosity="fct", alpha_limit=True, type_test=1) as s:
        s.set_state(z["u0"])
        ev = s.error_evaluator(z["jj"], z["rr"])
        t = float(z["time0"])
        worst = 0.0
        for tr, ratio in zip(z["times"], z["ratios"]):
            t, _ = s.run(1, time=t)
            assert t == tr
            worst = max(worst, abs(ev.ns_l1_ratio()[3] - ratio) / abs(ratio))
        ev.close()
    print("test 1, %d steps: largest relative gap of the unit-10 ratio %.3g" % (z["times"].size, worst), flush=True)
    assert worst <= 1e-11
