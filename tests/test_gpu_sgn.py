"""The SGN model of test 13 (5 components: h, hu, hv, h*eta, h*w) on the GPU, against the reference.

Fixture tests/golden/sgn_run.npz was written by the reference build on its own mesh
MESHES/SGN_CANAL/rectangle.3.FEM with its data (lambdaSGN = 9, CFL 0.08, h Dirichlet values on
sides 1 and 3 from the solitary wave of sol_anal, SOURCES/condlim.f90:751-827): euler in every
scheme defined for five components at two states, and 200 SSP-RK3 steps.  Everything is compared
bit for bit.  The per-stage Dirichlet values come from the host (SWE_BC_STAGES), evaluated by the
oracle's sol_anal at the stage times main.f90:78-90 uses.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO, golden, same_bits

pytestmark = pytest.mark.gpu

VISC = {"galerkin": "galerkin", "viscous": "viscous"}


def sgn_solver(z, viscosity="viscous", alpha=True, viscous_type=2):
    from swe_amd.solver import Solver
    g, htiny, hmax, cfl = z["scalars"]
    return Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], gravity=g, htiny=htiny, cfl=cfl,
                  max_water_h=hmax, viscosity=viscosity, viscous_type=viscous_type, alpha_limit=alpha, type_test=13,
                  lambda_sgn=float(z["lambda_sgn"]), h_dir=z["h_dir"], ux_dir=z["ux_dir"], uy_dir=z["uy_dir"])


def test_sgn_euler_every_scheme_bitwise():
    z = golden("sgn_run.npz")
    for k in range(2):
        for m, mode in enumerate(z["modes"]):
            visc, alpha, vt = str(mode).split("|")
            print("SGN euler", k, mode, flush=True)
            with sgn_solver(z, VISC[visc], bool(int(alpha)), 1 if vt == "type1" else 2) as s:
                assert s.compute_dt(z["states"][k]) == float(z["euler_dts"][k])
                un = s.euler(z["states"][k], float(z["euler_dts"][k]))
            assert same_bits(un, z["unext"][k, m]), (k, mode)


def test_sgn_run_against_reference(oracle):
    """200 steps, device-resident, with the solitary wave's h values per stage from the host."""
    from swe_amd.solver import BC_STAGES
    z = golden("sgn_run.npz")
    g, htiny = z["scalars"][0], z["scalars"][1]
    hd = z["h_dir"]
    xs, ys = z["rr"][hd, 0], z["rr"][hd, 1]

    def h_exact(t):
        return np.array([oracle.sol_anal(13, 1, x, y, t, g, htiny) for x, y in zip(xs, ys)])
    with sgn_solver(z) as s:
        s.set_bc(BC_STAGES)
        s.set_state(z["u0"])
        t = 0.0
        done = 0
        snaps = dict(zip([int(x) for x in z["snap_steps"]], z["snaps"]))
        for it in range(z["dts"].size):
            dt = s.step_begin()
            assert dt == z["dts"][it], it
            to = t
            t1 = t + dt
            t = to + dt
            t2 = t + dt / 2
            t = to + dt / 2
            t3 = t + dt
            s.step_finish(np.stack([h_exact(t1), h_exact(t2), h_exact(t3)]))
            t = to + dt
            if it == 0:
                assert same_bits(s.get_state(), z["s1"][2])
            if it + 1 in snaps:
                assert same_bits(s.get_state(), snaps[it + 1]), it + 1
            done = it + 1
        assert done == 200 and same_bits(s.get_state(), z["final"])
        assert s.time == float(z["time"]) and t == float(z["time"])


def test_sgn_batch_equals_runs_alone():
    """Two SGN meshes (no Dirichlet nodes) in one batch: each equals its solver run alone."""
    from swe_amd.batch import Batch
    from swe_amd.solver import assemble
    z = golden("sgn_run.npz")
    g, htiny, hmax, cfl = z["scalars"]
    prm = dict(gravity=g, htiny=htiny, cfl=cfl, max_water_h=hmax, viscosity="viscous", alpha_limit=True,
               type_test=13, lambda_sgn=9.0)
    from swe_amd import mesh as M
    m2 = M.renumber_first_touch(M.structured_mesh(61, 7, 0.0, 10.0, 0.0, 1.0, jitter=0.2, seed=4))
    cases = [(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], z["states"][1])]
    ia, ja, _d, _m, lumped, cij = assemble(m2.jj, m2.rr)
    u = np.zeros((m2.np, 5))
    x = m2.rr[:, 0]
    u[:, 0] = 0.1 + 0.01 * np.exp(-(x - 3.0) ** 2)
    u[:, 1] = 0.05 * u[:, 0]
    u[:, 3] = u[:, 0] * u[:, 0] * 1.01
    u[:, 4] = 0.001 * np.sin(x)
    cases.append((ia, ja, cij, lumped, np.zeros(m2.np), u))
    from swe_amd.solver import MeshArrays, Solver, make_params
    n = 30
    alone = []
    for (ia_, ja_, cij_, l_, b_, u_) in cases:
        with Solver(ia_, ja_, cij_, l_, b_, **prm) as s:
            s.set_state(u_)
            t, _ = s.run(n, time=0.0)
            alone.append((s.get_state(), t))
    rows = sum((c[0].size - 1 + 63) // 64 * 64 for c in cases) * 2
    with Batch(rows, int(sum(c[1].size for c in cases) * 1.5) + 4096, 64, 4) as b:
        ids = [b.add(MeshArrays(c[0], c[1], c[2], c[3], c[4]), make_params(**prm), c[5], max_steps=n) for c in cases]
        b.step(n + 2)
        for k, (u_ref, t_ref) in zip(ids, alone):
            st, t, steps, err = b.status(k)
            assert steps == n and err == 0 and t == t_ref
            assert same_bits(b.get_state(k), u_ref)


def test_sgn_refusals():
    """SWE_EBADOPT for what the reference leaves undefined with five components: entropy viscosity
    or FCT (Entprime(4:5,:) unset, update.f90:853-862) and Dirichlet lists for h*eta, h*w (sol_anal
    of those components after t = 1e-10, condlim.f90:801-826)."""
    from swe_amd.solver import SWE_EBADOPT, SweError
    z = golden("sgn_run.npz")
    for visc in ("entropy_visc", "fct"):
        with pytest.raises(SweError) as e:
            sgn_solver(z, visc)
        assert e.value.status == SWE_EBADOPT
    from swe_amd.solver import Solver
    g, htiny, hmax, cfl = z["scalars"]
    with pytest.raises(SweError) as e:
        Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], gravity=g, htiny=htiny, cfl=cfl, max_water_h=hmax,
               type_test=13, lambda_sgn=9.0, ux_dir=[0, 1])
    assert e.value.status == SWE_EBADOPT


REF_MAIN = os.path.join(REPO, "oracle", "_ref", "ref_main")
DROPIN_MAIN = os.path.join(REPO, "oracle", "_ref", "dropin_main")


@pytest.mark.skipif(not (os.path.exists(REF_MAIN) and os.path.exists(DROPIN_MAIN)),
                    reason="oracle/_ref was not built from the reference sources (make -C oracle ref)")
def test_reference_driver_runs_sgn_on_the_gpu(tmp_path):
    """SOURCES/main.f90 itself with test 13 on the reference's SGN canal mesh, linked once with the
    reference's update/fct and once with the MI355X binding: every printed step (time, dt) equal."""
    import pyoracle as O
    outs = []
    for exe in (REF_MAIN, DROPIN_MAIN):
        d = tmp_path / os.path.basename(exe)
        d.mkdir()
        with open(os.path.join(GOLDEN, "sgn_rectangle3.FEM"), "rb") as f, open(d / "mesh.FEM", "wb") as g:
            g.write(f.read())
        O.data_file(str(d / "data"), mesh_dir=".", mesh_file="mesh.FEM", test=13, tfinal=0.3, cfl=0.08,
                    viscosity="viscous", alpha=True, h_dir=(1, 3), lambda_sgn=9.0)
        print("running", exe, flush=True)
        r = subprocess.run([exe], cwd=str(d), capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        outs.append([l.strip() for l in r.stdout.splitlines() if l.strip().startswith(("time", "Min h"))])
    ref, gpu = outs
    assert len([l for l in ref if l.startswith("time")]) > 40
    assert gpu == ref


def test_sgn_canal_4m_nodes_against_oracle_and_properties(oracle):
    """The SGN model at scale: the solitary wave in the reference's canal on a 6401x641 lattice
    (4,103,041 nodes, viscous type2 + alpha): the first two steps bitwise against the oracle on the
    whole mesh, then 100 more steps through size-independent properties (every value finite, depth
    >= 0, h*eta > 0 where h > 0)."""
    from swe_amd import problems
    p = problems.sgn_canal()
    print("SGN canal", p.np, p.nnz, flush=True)
    pr = p.params
    orc = oracle.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=pr["gravity"], htiny=pr["htiny"], cfl=pr["cfl"],
                        max_water_h=pr["max_water_h"], viscosity_type=oracle.VISCOUS, viscous_type=2,
                        alpha_limit=True, type_test=13, lambda_sgn=pr["lambda_sgn"])
    try:
        ref, rdts, rt, err = orc.run(p.u0, 2)
    finally:
        orc.close()
    assert err == 0
    with p.solver() as s:
        s.set_state(p.u0)
        t, dts = s.run(2, time=0.0)
        u = s.get_state()
        assert u.shape == (p.np, 5) and same_bits(dts, rdts) and t == rt and same_bits(u, ref)
        t, dts = s.run(100)
        u = s.get_state()
        print("SGN canal after 102 steps: t = %.6g, min h %.6g" % (t, u[:, 0].min()), flush=True)
        assert np.all(np.isfinite(u)) and u[:, 0].min() >= 0.0 and np.all(dts > 0)
        assert np.all(u[u[:, 0] > 0, 3] > 0)
