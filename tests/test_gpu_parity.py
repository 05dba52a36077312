"""The HIP library (through its C-ABI) against the reference fixtures and the CPU checker.

Bar: bit-for-bit equality, except runs with Manning friction (pow), which must agree to
relative L1 and Linf <= 1e-11 in (h, qx, qy).
"""
import numpy as np
import pytest

from conftest import golden, rel_diff, same_bits

pytestmark = pytest.mark.gpu

TOL = 1e-11
RUNS = {
    "dambreak_run.npz": ("fct", True, 10, 0.02, 1),
    "dambreak9_run.npz": ("fct", True, 9, 0.02, 1),
    "bumps_run.npz": ("fct", True, 4, 0.0, 2),
    "cav025_run.npz": ("fct", True, 4, 0.0, 0),
    "cdisk_coriolis.npz": ("viscous", True, 12, 0.0, 0),
    "par_friction.npz": ("viscous", True, 11, 0.0, 0),
}


def set_lo_kind(s, kind):
    """Schedule kind of k_fct_lo: 0 lambda inline per row, 1 lambda entry-parallel in LDS."""
    s.schedule = {k: (b, r, kind if k == "k_fct_lo" else 0) for k, (b, r, _) in s.schedule.items()}


def solver_from(z, viscosity, alpha, type_test, mannings=0.0, viscous_type=2, bc=0, kind=0, roundoff_fix=True):
    from swe_amd.solver import Solver
    g, htiny, hmax, cfl = z["scalars"]
    s = Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], gravity=g, htiny=htiny, cfl=cfl,
               max_water_h=hmax, viscosity=viscosity, viscous_type=viscous_type, alpha_limit=alpha,
               type_test=type_test, mannings=mannings, h_dir=z["h_dir"], ux_dir=z["ux_dir"], uy_dir=z["uy_dir"],
               roundoff_fix=roundoff_fix)
    if bc == 2:
        amp, period, level = z["bc"]
        s.set_bc(2, amp, period, level)
    else:
        s.set_bc(bc)
    set_lo_kind(s, kind)
    return s


def test_compute_dij_bitwise():
    z = golden("par41_modes.npz")
    with solver_from(z, "viscous", False, 5) as s:
        for k in range(2):
            print("compute_dij state", k, flush=True)
            assert same_bits(s.compute_dij(z["states"][k]), z["dij_at_state"][k])


def test_lambda_bitwise_on_random_states(oracle):
    """compute_dij over a mesh with random wet/dry states: every |c_ij| lambda_ij bit for bit."""
    from swe_amd import problems
    p = problems.dam_break(61, 15, seed=5)
    rng = np.random.default_rng(1)
    orc = oracle.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0,
                        type_test=4)
    with p.solver(type_test=4) as s:
        for trial in range(5):
            u = np.zeros((p.np, 3))
            u[:, 0] = rng.uniform(0, 2, p.np) ** 3 * (rng.random(p.np) < 0.8)
            u[:, 1:] = rng.normal(0, 2.0, (p.np, 2)) * u[:, :1]
            d_ref, err = orc.compute_dij(u)
            assert err == 0
            print("random state", trial, flush=True)
            assert same_bits(s.compute_dij(u), d_ref)
            assert s.compute_dt(u) == orc.compute_dt(u)[0]


@pytest.mark.parametrize("kind", [0, 1])
def test_euler_every_mode_bitwise(kind):
    from swe_amd.solver import SWE_ENEG_H, SweError
    z = golden("par41_modes.npz")
    for k in range(2):
        for m, mode in enumerate(z["modes"]):
            visc, alpha, vt = str(mode).split("|")
            print("euler state", k, "mode", mode, "kind", kind, flush=True)
            with solver_from(z, visc, bool(int(alpha)), 5, viscous_type=1 if vt == "type1" else 2, kind=kind) as s:
                if z["status"][k, m] == 1:
                    with pytest.raises(SweError) as e:
                        s.euler(z["states"][k], float(z["dts"][k]))
                    assert e.value.status == SWE_ENEG_H
                else:
                    assert same_bits(s.euler(z["states"][k], float(z["dts"][k])), z["unext"][k, m]), mode


def test_fct_positivity_bitwise():
    from swe_amd.solver import Solver
    z = golden("fct.npz")
    npn = z["ia"].size - 1
    ia, ja = z["ia"], z["ja"]
    with Solver(ia, ja, np.ones((ja.size, 2)), np.ones(npn), np.zeros(npn), gravity=9.81, htiny=float(z["htiny"]),
                cfl=0.25, max_water_h=1.0) as s:
        assert same_bits(s.fct_positivity(z["us"], z["f"]), z["lij"])


def test_first_order_workload_2000_steps_bitwise():
    """Test 5 on the 81x81 grid, 2000 steps: the reference's final state and every dt."""
    from swe_amd import problems
    z = golden("par81_run.npz")
    p = problems.paraboloid(81)
    with p.solver() as s:
        s.set_state(p.u0)
        t, dts = s.run(1, time=0.0)
        assert same_bits(s.get_state(), z["s1"][2])
        s.set_state(p.u0)
        done, t = 0, 0.0
        for snap, step in zip(z["snaps"], z["snap_steps"]):
            print("paraboloid to step", step, flush=True)
            t, d = s.run(int(step) - done, time=t)
            assert same_bits(d, z["dts"][done:int(step)])
            assert same_bits(s.get_state(), snap), step
            done = int(step)
        assert same_bits(s.get_state(), z["final"])
        assert t == float(z["time"])


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("name", sorted(RUNS))
def test_runs_against_reference(name, kind):
    z = golden(name)
    visc, alpha, tt, mannings, bc = RUNS[name]
    with solver_from(z, visc, alpha, tt, mannings, bc=bc, kind=kind) as s:
        s.set_state(z["u0"])
        t, done = float(z["time0"]), 0
        for snap, step in zip(z["snaps"], z["snap_steps"]):
            print(name, "kind", kind, "to step", step, flush=True)
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
        l1, linf = rel_diff(u, z["final"])
        assert l1 <= TOL and linf <= TOL
        if mannings == 0:
            assert same_bits(u, z["final"]) and t == float(z["time"])
        assert u[:, 0].min() >= 0.0


def test_dam_break_63k_nodes_against_oracle(oracle):
    """Jittered 561x113 dam break (63,393 nodes), FCT + alpha + Manning, 30 steps."""
    from swe_amd import problems
    p = problems.dam_break(561, 113, seed=9)
    hd, uxd, uyd = p.dirichlet()
    orc = oracle.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0,
                        viscosity_type=oracle.FCT, alpha_limit=True, type_test=10, mannings=0.02)
    ref, rdts, rt, err = orc.run(p.u0, 30, bc_kind=1, nodes_ux=uxd, nodes_uy=uyd)
    assert err == 0
    with p.solver() as s:
        s.set_state(p.u0)
        t, d = s.run(30, time=0.0)
        u = s.get_state()
    l1, linf = rel_diff(u, ref)
    print("dam break 63k: rel L1 %.3g Linf %.3g, bitwise fraction %.6f" % (l1, linf, np.mean(u == ref)), flush=True)
    assert l1 <= TOL and linf <= TOL
    assert np.abs(d - rdts).max() <= TOL * rdts.max()


def test_bumps_inflow_66k_nodes_bitwise(oracle):
    """257x257 Delaunay mesh over Gaussian bumps with inflow, FCT + alpha, 20 steps."""
    from swe_amd import problems
    p = problems.bumps(n_side=257, L=1000.0, seed=4)
    hd, uxd, uyd = p.dirichlet()
    orc = oracle.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0,
                        viscosity_type=oracle.FCT, alpha_limit=True, type_test=4)
    kind, amp, period, level = p.bc
    ref, rdts, rt, err = orc.run(p.u0, 20, bc_kind=kind, amp=amp, period=period, level=level, nodes_h=hd,
                                 nodes_ux=uxd, nodes_uy=uyd)
    assert err == 0
    with p.solver() as s:
        s.set_state(p.u0)
        t, d = s.run(20, time=0.0)
        assert same_bits(d, rdts) and t == rt
        assert same_bits(s.get_state(), ref)


def test_negative_depth_is_reported():
    """check_hmin: the entropy-viscosity stage the reference stops on (Min h<0) returns SWE_ENEG_H."""
    from swe_amd.solver import SWE_ENEG_H, SweError
    z = golden("par41_modes.npz")
    assert z["status"][0, list(z["modes"]).index("entropy_visc|1|type2")] == 1
    with solver_from(z, "entropy_visc", True, 5) as s:
        s.set_state(z["u0"])
        with pytest.raises(SweError) as e:
            s.run(1, time=0.0)
        assert e.value.status == SWE_ENEG_H


def test_per_stage_boundary_values_against_oracle(oracle):
    """Device-resident steps whose Dirichlet values the host evaluates per stage once dt is known
    (swe_step_begin / swe_step_finish; bdy of SOURCES/main.f90:176-188 with a time-dependent
    boundary function), against the CPU checker driven through the same time loop."""
    from swe_amd import problems
    from swe_amd.solver import BC_STAGES
    p = problems.bumps(n_side=65, L=500.0, seed=5)
    hd, uxd, uyd = (np.asarray(x, np.int64) for x in p.dirichlet())
    z = p.bath
    nh, nx = hd.size, uxd.size

    def values(t):
        h = np.maximum(1.0 - z[hd] + 0.05 * np.sin(2 * np.pi * t / 7.0), 0.0)
        return np.concatenate([h, np.zeros(uxd.size), np.zeros(uyd.size)])

    def bdy(u, v):
        u[hd, 0] = v[:nh]
        u[uxd, 1] = v[nh:nh + nx]
        u[uyd, 2] = v[nh + nx:]
        return u

    def stage_times(to, dt):
        return [to + dt, (to + dt) + dt / 2, (to + dt / 2) + dt]

    nsteps = 30
    orc = oracle.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0,
                        viscosity_type=oracle.FCT, alpha_limit=True, type_test=4)
    u, t, dts = p.u0.copy(), 0.0, []
    for _ in range(nsteps):
        dt, err = orc.compute_dt(u)
        assert err == 0
        ts = stage_times(t, dt)
        uo = u
        u1, _, e1 = orc.euler(uo, dt)
        u1 = bdy(u1, values(ts[0]))
        ui, _, e2 = orc.euler(u1, dt)
        u2 = bdy((3 * uo + ui) / 4, values(ts[1]))
        ui, _, e3 = orc.euler(u2, dt)
        u = bdy((uo + 2 * ui) / 3, values(ts[2]))
        assert e1 == e2 == e3 == 0
        t = t + dt
        dts.append(dt)
    with p.solver() as s:
        s.set_state(p.u0)
        s.set_bc(BC_STAGES)
        s.time = 0.0
        td = 0.0
        for k in range(nsteps):
            dt = s.step_begin()
            assert dt == dts[k], k
            s.step_finish(np.stack([values(x) for x in stage_times(td, dt)]))
            td = td + dt
        assert s.time == t == td
        assert same_bits(s.get_state(), u)


def test_scrambled_numbering_against_oracle(oracle):
    """A mesh numbered at random (neighbours up to the whole mesh apart, so many column offsets do
    not fit the 16-bit form and are read from the 32-bit columns), FCT + alpha with inflow, 10 steps,
    bit for bit against the CPU checker on the same numbering."""
    from swe_amd import problems
    from swe_amd.solver import Solver, assemble
    p = problems.bumps(n_side=201, L=1000.0, seed=8)
    perm = np.random.default_rng(1).permutation(p.np)          # new number of node k
    jj = perm[p.mesh.jj]
    rr = np.empty_like(p.mesh.rr)
    rr[perm] = p.mesh.rr
    ia, ja, diag, mass, lumped, cij = assemble(jj, rr)
    assert np.abs(ja - np.repeat(np.arange(p.np), np.diff(ia))).max() > 32767
    bath = np.empty_like(p.bath)
    bath[perm] = p.bath
    u0 = np.empty_like(p.u0)
    u0[perm] = p.u0
    hd, uxd, uyd = (np.sort(perm[np.asarray(x, np.int64)]) for x in p.dirichlet())
    kind, amp, period, level = p.bc
    orc = oracle.Oracle(ia, ja, cij, lumped, bath, gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0,
                        viscosity_type=oracle.FCT, alpha_limit=True, type_test=4)
    ref, rdts, rt, err = orc.run(u0, 10, bc_kind=kind, amp=amp, period=period, level=level, nodes_h=hd,
                                 nodes_ux=uxd, nodes_uy=uyd)
    assert err == 0
    with Solver(ia, ja, cij, lumped, bath, h_dir=hd, ux_dir=uxd, uy_dir=uyd, **p.params) as s:
        s.set_bc(kind, amp, period, level)
        s.set_state(u0)
        t, d = s.run(10, time=0.0)
        assert same_bits(d, rdts) and t == rt
        assert same_bits(s.get_state(), ref)


def test_fct_generic_against_oracle(oracle):
    """FCT_generic (SAFE/fct.f90:3-43) on the GPU, bit for bit against the CPU checker."""
    from swe_amd.solver import fct_generic_csr
    rng = np.random.default_rng(11)
    for n, k in ((50, 4), (5000, 7), (60000, 9)):
        rows = [np.unique(np.concatenate([[i], rng.integers(0, n, k)])) for i in range(n)]
        ia = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
        ja = np.concatenate(rows).astype(np.int32)
        f = rng.normal(size=ja.size)
        u = rng.uniform(0, 1, n)
        maxn = u + rng.uniform(0, 0.05, n)
        minn = u - rng.uniform(0, 0.05, n)
        dg = rng.uniform(0.5, 2.0, n)
        print("fct_generic", n, flush=True)
        assert same_bits(fct_generic_csr(ia, ja, u, maxn, minn, f, dg), oracle.fct_generic(ia, ja, f, u, maxn, minn, dg))


@pytest.mark.parametrize("kind", [0, 1])
def test_roundoff_off_branches_against_reference(kind):
    """The reference built with if_roundoff_fix = .FALSE. (tests/golden/noround.npz: smb_2 and the
    clipped update): euler in five modes at two states, 150 first-order and 100 fct + alpha steps,
    every bit, with both schedule kinds of k_fct_lo."""
    z = golden("par41_modes.npz")
    n = golden("noround.npz")
    for k in range(2):
        for m, mode in enumerate(n["modes"]):
            visc, alpha, _ = str(mode).split("|")
            print("roundoff off: euler state", k, "mode", mode, "kind", kind, flush=True)
            with solver_from(z, visc, bool(int(alpha)), 5, kind=kind, roundoff_fix=False) as s:
                assert same_bits(s.euler(z["states"][k], float(z["dts"][k])), n["unext"][k, m]), mode
    for tag, visc, alpha, steps in (("viscous", "viscous", False, 150), ("fct_alpha", "fct", True, 100)):
        print("roundoff off: run", tag, flush=True)
        with solver_from(z, visc, alpha, 5, kind=kind, roundoff_fix=False) as s:
            s.set_state(z["u0"])
            t, dts = s.run(steps, time=float(z["time0"]))
            assert same_bits(dts, n["run_%s_dts" % tag]) and t == float(n["run_%s_time" % tag])
            assert same_bits(s.get_state(), n["run_%s_final" % tag]), tag
