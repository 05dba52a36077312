"""Write the golden fixtures of tests/golden/ from the reference Fortran built from source.

Run in the container that holds /root/reference, after `make -C oracle ref`:

    python tests/golden/make_golden.py

Every fixture is data: the inputs given to the reference driver (oracle/ref_driver.f90, which
links the reference modules) and the outputs it wrote.  Cases:

  lambda.npz          compute_lambda_vacc (SOURCES/condlim.f90:879-957) on random and edge-case states
  fct.npz             FCT_positivity (SOURCES/fct.f90:5-42) on random fluxes over a mesh pattern
  par41_modes.npz     euler (SOURCES/update.f90:134-289) for every viscosity mode, test 5 on 41x41
  par81_run.npz       2000 SSP-RK3 steps of test 5 on the 81x81 grid (the first-order workload)
  dambreak_run.npz    FCT + alpha + Manning (type_test 10) dam break, jittered 101x21, walls
  dambreak9_run.npz   the same scheme with test 9's fixed htiny and dt (SOURCES/main.f90:66-69), 51x11
  sgn_run.npz         the SGN model of test 13 (5 components, mSGN_RHS) on MESHES/SGN_CANAL/rectangle.3.FEM
  bumps_run.npz       FCT + alpha, Gaussian bumps, sinusoidal inflow, 65x65 Delaunay mesh
  cav025_run.npz      the reference's own mesh MESHES/DAM_BREAK/cav_0.25.FEM, test 4 init, FCT + alpha
  cdisk_coriolis.npz  the reference's own mesh MESHES/PARABOLOID/cdisk.1_mesh.FEM, test 12 (Coriolis)
  par_friction.npz    test 11 (linear friction) on a 41x41 grid of the paraboloid, viscous + alpha
  errors.npz          the quadrature norms of compute_errors (SOURCES/main.f90:198-263) on states of the
                      fixtures above (tests 4, 5, 11, 12), by the reference's sol_anal, ns_anal_l1, ns_anal_0
  errors_bump.npz     the same norms for the steady bump flows (tests 3, 6, 7) on a [0,25]x[0,1] strip
  flow.npz            the time loop's unit-10 diagnostic (SOURCES/main.f90:94-99, ns_l1 of qx, qy, h*sqrt(h))
                      on the dam-break states (test 10) and the SGN canal's final state
  unit10_test1.npz    the reference program (main.f90) on test 1, FCT + alpha: set-up arrays, and per step
                      the time and the ratio it wrote to fort.10
  noround.npz         the reference built with if_roundoff_fix = .FALSE. (its smb_2 branches): euler in five
                      modes at the par41_modes.npz states, 150 first-order and 100 fct + alpha steps
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "shallow-water-eric_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

from swe_amd import mesh as M  # noqa: E402
from swe_amd import problems as P  # noqa: E402
import pyoracle as O  # noqa: E402

REF = "/root/reference"
SEED = 20261202


def save(name, **arrays):
    np.savez_compressed(os.path.join(HERE, name), **arrays)
    print("wrote", name, "%.1f KB" % (os.path.getsize(os.path.join(HERE, name)) / 1024))


def fresh(d, sub):
    p = os.path.join(d, sub)
    os.makedirs(p, exist_ok=True)
    return p


def read_state(path, npn, nc=3):
    return np.fromfile(path, dtype="<f8").reshape(npn, nc)


# ----------------------------------------------------------------------------------------------
def gen_lambda():
    rng = np.random.default_rng(SEED)
    g, htiny = 9.81, 4.0e-4
    recs = []
    n = 4000
    th = rng.uniform(0, 2 * np.pi, n)
    hl = rng.uniform(0, 2, n) ** 2
    hr = rng.uniform(0, 2, n) ** 2
    ql = rng.normal(0, 1.5, (n, 2)) * hl[:, None]
    qr = rng.normal(0, 1.5, (n, 2)) * hr[:, None]
    recs.append(np.column_stack([hl, ql, hr, qr, np.cos(th), np.sin(th)]))
    # wet/dry and near-threshold edge cases
    m = 1000
    th = rng.uniform(0, 2 * np.pi, m)
    choices = np.array([0.0, htiny * 0.5, htiny, htiny * 1.0000001, 1e-3, 0.1, 1.0])
    hl = rng.choice(choices, m)
    hr = rng.choice(choices, m)
    ql = rng.normal(0, 0.5, (m, 2)) * hl[:, None]
    qr = rng.normal(0, 0.5, (m, 2)) * hr[:, None]
    recs.append(np.column_stack([hl, ql, hr, qr, np.cos(th), np.sin(th)]))
    # strong rarefactions (case 3) and strong shocks
    k = 1000
    th = rng.uniform(0, 2 * np.pi, k)
    hl = rng.uniform(0.01, 1, k)
    hr = rng.uniform(0.01, 1, k)
    vv = rng.uniform(-30, 30, k)
    ql = np.column_stack([-vv * np.cos(th), -vv * np.sin(th)]) * hl[:, None]
    qr = np.column_stack([vv * np.cos(th), vv * np.sin(th)]) * hr[:, None]
    recs.append(np.column_stack([hl, ql, hr, qr, np.cos(th), np.sin(th)]))
    # both dry, equal states
    recs.append(np.array([[0, 0, 0, 0, 0, 0, 1, 0], [0.5, 0.1, -0.2, 0.5, 0.1, -0.2, 0.6, 0.8]], dtype=float))
    rec = np.ascontiguousarray(np.vstack(recs))
    d = O.tmpdir()
    with open(os.path.join(d, "in.bin"), "wb") as f:
        f.write(np.array([rec.shape[0], 5], dtype="<i4").tobytes())
        f.write(np.array([g, htiny], dtype="<f8").tobytes())
        f.write(rec.astype("<f8").tobytes())
    O.run_ref(["lambda", "in.bin", "out.bin"], cwd=d)
    lam = np.fromfile(os.path.join(d, "out.bin"), dtype="<f8")
    shutil.rmtree(d)
    save("lambda.npz", rec=rec, lam=lam, gravity=g, htiny=htiny)


def gen_fct():
    rng = np.random.default_rng(SEED + 1)
    m = M.renumber_first_touch(M.structured_mesh(21, 17, 0.0, 1.0, 0.0, 1.0, jitter=0.2, seed=5))
    ia, ja, *_ = O.assemble(m.jj, m.rr)
    npn, nnz = ia.size - 1, ja.size
    f = rng.normal(0, 1.0, nnz)
    f[rng.random(nnz) < 0.1] = 0.0
    us = rng.uniform(-1e-4, 2.0, npn)
    us[rng.random(npn) < 0.1] = 0.0
    htiny = 1e-3
    d = O.tmpdir()
    with open(os.path.join(d, "in.bin"), "wb") as fh:
        fh.write(np.array([npn, nnz], dtype="<i4").tobytes())
        fh.write(np.array([htiny], dtype="<f8").tobytes())
        fh.write((ia + 1).astype("<i4").tobytes())
        fh.write((ja + 1).astype("<i4").tobytes())
        fh.write(f.astype("<f8").tobytes())
        fh.write(us.astype("<f8").tobytes())
    O.run_ref(["fct", "in.bin", "out.bin"], cwd=d)
    lij = np.fromfile(os.path.join(d, "out.bin"), dtype="<f8")
    shutil.rmtree(d)
    save("fct.npz", ia=ia, ja=ja, f=f, us=us, htiny=htiny, lij=lij)


# ----------------------------------------------------------------------------------------------
class RefCase:
    """A mesh written to a scratch directory with a reference `data` file (and an init file)."""

    def __init__(self, raw_mesh=None, fem_path=None, *, test, viscosity="viscous", alpha=False, cfl=0.25,
                 viscous_type="type2", h_sides=(), ux_sides=(), uy_sides=(), mannings=None,
                 init=None, bc=(0, 0.0, 1.0, 0.0), lambda_sgn=None, driver=None):
        self.nc = 5 if test == 13 else 3
        self.driver = driver                  # None: the unmodified reference build
        self.d = O.tmpdir()
        if fem_path:
            shutil.copy(fem_path, os.path.join(self.d, "mesh.FEM"))
        else:
            M.write_fem(os.path.join(self.d, "mesh.FEM"), raw_mesh)
        self.kw = dict(mesh_dir=".", mesh_file="mesh.FEM", test=test, cfl=cfl, viscous_type=viscous_type,
                       h_dir=h_sides, ux_dir=ux_sides, uy_dir=uy_sides, mannings=mannings, lambda_sgn=lambda_sgn)
        self.write_data(viscosity, alpha)
        self.init = "none"
        if init is not None:
            bath, u0, g, hmax = init
            O.write_custom_init(os.path.join(self.d, "init.bin"), bath, u0, gravity=g, max_water_h=hmax,
                                bc_kind=bc[0], amp=bc[1], period=bc[2], level=bc[3])
            self.init = "init.bin"

    def write_data(self, viscosity, alpha, viscous_type=None):
        kw = dict(self.kw)
        if viscous_type:
            kw["viscous_type"] = viscous_type
        O.data_file(os.path.join(self.d, "data"), viscosity=viscosity, alpha=alpha, **kw)

    def setup(self):
        out = fresh(self.d, "setup")
        O.run_ref(["setup", "data", self.init, "setup"], cwd=self.d, driver=self.driver)
        return O.read_setup(out)

    def euler(self, u, dt):
        """Returns (unext, dij, status): status 1 when the reference stopped in check_hmin."""
        np.ascontiguousarray(u, dtype="<f8").tofile(os.path.join(self.d, "state.f8"))
        out = fresh(self.d, "eu")
        for f in os.listdir(out):
            os.remove(os.path.join(out, f))
        try:
            O.run_ref(["euler", "data", self.init, "state.f8", repr(float(dt)), "eu"], cwd=self.d, driver=self.driver)
        except RuntimeError as e:
            if "Min h<0, STOP" not in str(e):
                raise
            return np.full(u.shape, np.nan), None, 1
        return (read_state(os.path.join(out, "unext.f8"), u.shape[0], self.nc), np.fromfile(os.path.join(out, "dij.f8")),
                0)

    def dij(self, u):
        np.ascontiguousarray(u, dtype="<f8").tofile(os.path.join(self.d, "state.f8"))
        O.run_ref(["dij", "data", self.init, "state.f8", "dij.f8"], cwd=self.d, driver=self.driver)
        return np.fromfile(os.path.join(self.d, "dij.f8"))

    def run(self, nsteps, every=0):
        out = fresh(self.d, "run")
        args = ["run", "data", self.init, nsteps, "run"] + ([every] if every else [])
        O.run_ref(args, cwd=self.d, timeout=7200, driver=self.driver)
        npn = int(np.fromfile(os.path.join(self.d, "setup", "dims.i4"), dtype="<i4")[0]) if os.path.exists(
            os.path.join(self.d, "setup", "dims.i4")) else None
        u0 = np.fromfile(os.path.join(out, "u0.f8"))
        nc = self.nc
        npn = u0.size // nc
        res = dict(final=read_state(os.path.join(out, "final.f8"), npn, nc),
                   dts=np.fromfile(os.path.join(out, "dt.f8")),
                   time=np.fromfile(os.path.join(out, "time.f8"))[0],
                   s1=np.stack([read_state(os.path.join(out, "s1_stage%d.f8" % k), npn, nc) for k in (1, 2, 3)]))
        snaps = sorted(f for f in os.listdir(out) if f.startswith("snap_"))
        if snaps:
            res["snaps"] = np.stack([read_state(os.path.join(out, f), npn, nc) for f in snaps])
            res["snap_steps"] = np.array([int(f[5:11]) for f in snaps])
        return res

    def close(self):
        shutil.rmtree(self.d, ignore_errors=True)


def setup_arrays(S):
    return dict(rr=S["rr"], jj=S["jj"], ia=S["ia"], ja=S["ja"], lumped=S["lumped"], cij=S["cij"],
                bath=S["bath"], u0=S["u0"], h_dir=S["h_js_D"], ux_dir=S["ux_js_D"],
                uy_dir=S["uy_js_D"], scalars=np.array([S["gravity"], S["htiny"], S["max_water_h"], S["cfl"]]),
                type_test=S["type_test"], time0=S["time0"])


def gen_par41_modes():
    raw = M.structured_mesh(41, 41, 0.0, 4.0, 0.0, 4.0)
    case = RefCase(raw, test=5, viscosity="viscous", cfl=0.25)
    S = case.setup()
    # a later state with wet/dry fronts: 150 first-order steps
    r = case.run(150)
    states = np.stack([S["u0"], r["final"]])
    orc = O.Oracle(S["ia"], S["ja"], S["cij"], S["lumped"], S["bath"], gravity=S["gravity"], htiny=S["htiny"],
                   cfl=S["cfl"], max_water_h=S["max_water_h"], type_test=5)
    modes = [("viscous", False, "type2"), ("viscous", True, "type2"), ("fct", True, "type2"), ("fct", False, "type2"),
             ("entropy_visc", True, "type2"), ("galerkin", False, "type2"), ("viscous", True, "type1")]
    outs, dijs, dts, dij0, status = [], [], [], [], []
    for u in states:
        dt, _ = orc.compute_dt(u)
        dts.append(dt)
        dij0.append(case.dij(u))
        for (v, a, t) in modes:
            case.write_data(v, a, t)
            un, dij, st = case.euler(u, dt)
            outs.append(un)
            dijs.append(dij if dij is not None else np.full(S["ja"].size, np.nan))
            status.append(st)
        case.write_data("viscous", False, "type2")
    save("par41_modes.npz", **setup_arrays(S), states=states, dts=np.array(dts), dij_at_state=np.stack(dij0),
         modes=np.array(["%s|%d|%s" % m for m in modes]), unext=np.stack(outs).reshape(2, len(modes), -1, 3),
         dij_after=np.stack(dijs).reshape(2, len(modes), -1), status=np.array(status).reshape(2, len(modes)))
    case.close()


def gen_par81_run():
    raw = M.structured_mesh(81, 81, 0.0, 4.0, 0.0, 4.0)
    case = RefCase(raw, test=5, viscosity="viscous", cfl=0.25)
    S = case.setup()
    r = case.run(2000, every=500)
    save("par81_run.npz", u0=S["u0"], lumped=S["lumped"], final=r["final"], dts=r["dts"], time=r["time"],
         s1=r["s1"], snaps=r["snaps"], snap_steps=r["snap_steps"])
    case.close()


def custom_case(problem, *, nsteps, every=0, test, mannings=None, viscosity="fct", alpha=True, raw=None):
    case = RefCase(raw, test=test, viscosity=viscosity, alpha=alpha, cfl=problem.params["cfl"],
                   h_sides=problem.h_sides, ux_sides=problem.ux_sides, uy_sides=problem.uy_sides, mannings=mannings,
                   init=(problem.bath, problem.u0, problem.params["gravity"], problem.params["max_water_h"]),
                   bc=problem.bc)
    S = case.setup()
    r = case.run(nsteps, every=every)
    return case, S, r


def gen_dambreak():
    p = P.dam_break(101, 21, seed=7)
    raw = M.structured_mesh(101, 21, 0.0, 100.0, 0.0, 20.0, jitter=0.2, seed=7)
    case, S, r = custom_case(p, nsteps=200, every=50, test=10, mannings=0.02, raw=raw)
    dt, _ = O.Oracle(S["ia"], S["ja"], S["cij"], S["lumped"], S["bath"], gravity=9.81, htiny=S["htiny"],
                     cfl=S["cfl"], max_water_h=1.0, type_test=10, mannings=0.02).compute_dt(r["snaps"][0])
    eu, dij, st = case.euler(r["snaps"][0], dt)
    assert st == 0
    save("dambreak_run.npz", **setup_arrays(S), final=r["final"], dts=r["dts"], time=r["time"], s1=r["s1"],
         snaps=r["snaps"], snap_steps=r["snap_steps"], euler_in=r["snaps"][0], euler_dt=dt, euler_out=eu,
         euler_dij=dij)
    case.close()


def gen_dambreak9():
    """Test 9's branch of the time loop (SOURCES/main.f90:66-69: htiny = 1e-3, dt = 2.5e-2, no
    COMPUTE_DT) with its Manning friction, on a coarse dam break (test 9's own Malpasset mesh is not
    in the reference tree)."""
    p = P.dam_break(51, 11, seed=9)
    raw = M.structured_mesh(51, 11, 0.0, 100.0, 0.0, 20.0, jitter=0.2, seed=9)
    case, S, r = custom_case(p, nsteps=120, every=40, test=9, mannings=0.02, raw=raw)
    save("dambreak9_run.npz", **setup_arrays(S), final=r["final"], dts=r["dts"], time=r["time"], s1=r["s1"],
         snaps=r["snaps"], snap_steps=r["snap_steps"])
    case.close()


def gen_sgn():
    """The SGN model (test 13, 5 components: h, hu, hv, h*eta, h*w) on the reference's own mesh
    MESHES/SGN_CANAL/rectangle.3.FEM with its data (lambdaSGN = 9, CFL 0.08, h Dirichlet on sides
    1 and 3 from the solitary-wave sol_anal): 200 steps of the first-order scheme with the alpha
    limiter, then euler in the schemes defined for 5 components (viscous with and without alpha,
    type1, galerkin) at the initial and the final state."""
    fem = os.path.join(REF, "MESHES", "SGN_CANAL", "rectangle.3.FEM")
    case = RefCase(fem_path=fem, test=13, viscosity="viscous", alpha=True, cfl=0.08, h_sides=(1, 3),
                   lambda_sgn=9.0)
    S = case.setup()
    r = case.run(200, every=50)
    states = np.stack([S["u0"], r["final"]])
    orc = O.Oracle(S["ia"], S["ja"], S["cij"], S["lumped"], S["bath"], gravity=S["gravity"], htiny=S["htiny"],
                   cfl=S["cfl"], max_water_h=S["max_water_h"], type_test=13, lambda_sgn=9.0)
    modes = [("viscous", True, "type2"), ("viscous", False, "type2"), ("viscous", True, "type1"),
             ("galerkin", False, "type2")]
    outs, dts = [], []
    for u in states:
        dt, _ = orc.compute_dt(u)
        dts.append(dt)
        for (v, a, t) in modes:
            case.write_data(v, a, t)
            un, dij, st = case.euler(u, dt)
            assert st == 0
            outs.append(un)
        case.write_data("viscous", True, "type2")
    save("sgn_run.npz", **setup_arrays(S), final=r["final"], dts=r["dts"], time=r["time"], s1=r["s1"],
         snaps=r["snaps"], snap_steps=r["snap_steps"], lambda_sgn=9.0, states=states, euler_dts=np.array(dts),
         modes=np.array(["%s|%d|%s" % m for m in modes]), unext=np.stack(outs).reshape(2, len(modes), -1, 5))
    case.close()


def gen_bumps():
    p = P.bumps(n_side=65, L=1000.0, seed=11)
    raw = M.jittered_delaunay_mesh(65, 1000.0, jitter=0.25, seed=11)
    case, S, r = custom_case(p, nsteps=200, every=50, test=4, raw=raw)
    save("bumps_run.npz", **setup_arrays(S), final=r["final"], dts=r["dts"], time=r["time"], s1=r["s1"],
         snaps=r["snaps"], snap_steps=r["snap_steps"], bc=np.array(p.bc[1:], dtype=float))
    case.close()


def gen_cav025():
    fem = os.path.join(REF, "MESHES", "DAM_BREAK", "cav_0.25.FEM")
    case = RefCase(fem_path=fem, test=4, viscosity="fct", alpha=True, cfl=0.25)
    S = case.setup()
    r = case.run(100, every=25)
    save("cav025_run.npz", **setup_arrays(S), final=r["final"], dts=r["dts"], time=r["time"], s1=r["s1"],
         snaps=r["snaps"], snap_steps=r["snap_steps"])
    case.close()


def gen_cdisk_coriolis():
    fem = os.path.join(REF, "MESHES", "PARABOLOID", "cdisk.1_mesh.FEM")
    case = RefCase(fem_path=fem, test=12, viscosity="viscous", alpha=True, cfl=0.3)
    S = case.setup()
    r = case.run(100, every=25)
    save("cdisk_coriolis.npz", **setup_arrays(S), final=r["final"], dts=r["dts"], time=r["time"], s1=r["s1"],
         snaps=r["snaps"], snap_steps=r["snap_steps"])
    case.close()


def gen_par_friction():
    raw = M.structured_mesh(41, 41, 0.0, 4.0, 0.0, 4.0)
    case = RefCase(raw, test=11, viscosity="viscous", alpha=True, cfl=0.25)
    S = case.setup()
    r = case.run(100, every=25)
    save("par_friction.npz", **setup_arrays(S), final=r["final"], dts=r["dts"], time=r["time"], s1=r["s1"],
         snaps=r["snaps"], snap_steps=r["snap_steps"])
    case.close()


def gen_errors():
    """compute_errors' sums, (3, 4) per case: [L1 error, L1 norm, L2 error, L2 norm] of h, qx, qy.
    The states and meshes are those of the fixtures named in each key."""
    out = {}

    def errs(case, u, t):
        np.ascontiguousarray(u, dtype="<f8").tofile(os.path.join(case.d, "state.f8"))
        O.run_ref(["errors", "data", case.init, "state.f8", repr(float(t)), "errors.f8"], cwd=case.d)
        return np.fromfile(os.path.join(case.d, "errors.f8")).reshape(3, 4)

    z = np.load(os.path.join(HERE, "par41_modes.npz"))
    case = RefCase(M.structured_mesh(41, 41, 0.0, 4.0, 0.0, 4.0), test=5, viscosity="viscous", cfl=0.25)
    S = case.setup()
    assert np.array_equal(S["jj"], z["jj"]) and np.array_equal(S["rr"], z["rr"])
    out["par41_modes|states0|0.0"] = errs(case, z["states"][0], 0.0)
    out["par41_modes|states1|0.75"] = errs(case, z["states"][1], 0.75)
    case.close()
    for name, fem, test in (("cav025_run", os.path.join(REF, "MESHES", "DAM_BREAK", "cav_0.25.FEM"), 4),
                            ("cdisk_coriolis", os.path.join(REF, "MESHES", "PARABOLOID", "cdisk.1_mesh.FEM"), 12),
                            ("par_friction", None, 11)):
        z = np.load(os.path.join(HERE, name + ".npz"))
        if fem:
            case = RefCase(fem_path=fem, test=test, viscosity="fct", alpha=True)
        else:
            case = RefCase(M.structured_mesh(41, 41, 0.0, 4.0, 0.0, 4.0), test=test, viscosity="viscous", alpha=True)
        S = case.setup()
        assert np.array_equal(S["jj"], z["jj"]) and np.array_equal(S["rr"], z["rr"])
        out["%s|final|time" % name] = errs(case, z["final"], float(z["time"]))
        case.close()
    # the SGN model (test 13: h, qx, qy of its five components) on the canal of sgn_run.npz, at the
    # initial state and after the 200 steps
    z = np.load(os.path.join(HERE, "sgn_run.npz"))
    case = RefCase(fem_path=os.path.join(REF, "MESHES", "SGN_CANAL", "rectangle.3.FEM"), test=13, viscosity="viscous",
                   alpha=True, cfl=0.08, h_sides=(1, 3), lambda_sgn=9.0)
    S = case.setup()
    assert np.array_equal(S["jj"], z["jj"]) and np.array_equal(S["rr"], z["rr"])
    out["sgn_run|states0|0.0"] = errs(case, z["states"][0], 0.0)
    out["sgn_run|final|time"] = errs(case, z["final"], float(z["time"]))
    case.close()
    save("errors.npz", **out)
    # the steady flows over a bump (tests 3, 6, 7) on a strip of [0,25]x[0,1], at their initial states,
    # at t = 0 and at t = 1 (the analytic steady state)
    raw = M.structured_mesh(101, 5, 0.0, 25.0, 0.0, 1.0)
    bump = {}
    for test in (3, 6, 7):
        case = RefCase(raw, test=test, viscosity="viscous")
        S = case.setup()
        bump["jj"], bump["rr"] = S["jj"], S["rr"]
        bump["u0_%d" % test] = S["u0"]
        bump["scalars_%d" % test] = np.array([S["gravity"], S["htiny"]])
        for t in (0.0, 1.0):
            bump["errors_%d_t%d" % (test, int(t))] = errs(case, S["u0"], t)
        case.close()
    save("errors_bump.npz", **bump)


def gen_flow():
    """The time loop's unit-10 diagnostic for tests 1, 2 and 10 (main.f90:94-99: ns_l1 of qx, qy and
    h*sqrt(h), and 0.5*(q1+q2)/(g*q3)), 4 reals per key, on the states of the dam-break fixture (test
    10, the case that writes it) and on the final state of the SGN canal (five components per node)."""
    out = {}

    def flow(case, u):
        np.ascontiguousarray(u, dtype="<f8").tofile(os.path.join(case.d, "state.f8"))
        O.run_ref(["flow", "data", case.init, "state.f8", "flow.f8"], cwd=case.d)
        return np.fromfile(os.path.join(case.d, "flow.f8"))

    z = np.load(os.path.join(HERE, "dambreak_run.npz"))
    p = P.dam_break(101, 21, seed=7)
    raw = M.structured_mesh(101, 21, 0.0, 100.0, 0.0, 20.0, jitter=0.2, seed=7)
    case = RefCase(raw, test=10, viscosity="fct", alpha=True, cfl=p.params["cfl"], h_sides=p.h_sides,
                   ux_sides=p.ux_sides, uy_sides=p.uy_sides, mannings=0.02,
                   init=(p.bath, p.u0, p.params["gravity"], p.params["max_water_h"]), bc=p.bc)
    S = case.setup()
    assert np.array_equal(S["jj"], z["jj"]) and np.array_equal(S["rr"], z["rr"])
    out["dambreak_run|u0"] = flow(case, z["u0"])
    for k, st in enumerate(z["snap_steps"]):
        out["dambreak_run|snap%d" % int(st)] = flow(case, z["snaps"][k])   # the last snapshot is the final state
    case.close()
    z = np.load(os.path.join(HERE, "sgn_run.npz"))
    case = RefCase(fem_path=os.path.join(REF, "MESHES", "SGN_CANAL", "rectangle.3.FEM"), test=13, viscosity="viscous",
                   alpha=True, cfl=0.08, h_sides=(1, 3), lambda_sgn=9.0)
    S = case.setup()
    assert np.array_equal(S["jj"], z["jj"]) and np.array_equal(S["rr"], z["rr"])
    out["sgn_run|final"] = flow(case, z["final"])
    case.close()
    save("flow.npz", **out)
    # the reference program itself (SOURCES/main.f90, oracle/_ref/ref_main) on test 1, FCT + alpha, to
    # Tfinal = 1: its set-up arrays and, per step, the time and the ratio it wrote to unit 10 (fort.10)
    case = RefCase(M.structured_mesh(81, 41, 0.0, 2.0, 0.0, 1.0, jitter=0.2, seed=3), test=1, viscosity="fct",
                   alpha=True, cfl=0.25)
    S = case.setup()
    r = subprocess.run([os.path.join(REPO, "oracle", "_ref", "ref_main")], cwd=case.d, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:]
    rows = np.array([[float(x) for x in l.split()] for l in open(os.path.join(case.d, "fort.10")).read().splitlines()])
    save("unit10_test1.npz", **setup_arrays(S), times=rows[:, 0], ratios=rows[:, 1])
    case.close()


def gen_noround():
    """The reference built with if_roundoff_fix = .FALSE. (MODULE update's compile-time switch,
    SOURCES/update.f90:15; oracle/Makefile target ref_driver_noround): euler in the five modes whose
    branches it changes (smb_2 and the clipped update max(h + dt*rk, 0) instead of smb_2_roundoff and
    the fix_roundoff update, update.f90:193-215, 246-262) at the two states of par41_modes.npz, and
    runs of 150 first-order and 100 fct + alpha steps from the initial state."""
    z = np.load(os.path.join(HERE, "par41_modes.npz"))
    case = RefCase(M.structured_mesh(41, 41, 0.0, 4.0, 0.0, 4.0), test=5, viscosity="viscous", cfl=0.25,
                   driver=O.REF_DRIVER_NOROUND)
    S = case.setup()
    assert np.array_equal(S["jj"], z["jj"]) and np.array_equal(S["rr"], z["rr"])
    modes = [("viscous", False, "type2"), ("viscous", True, "type2"), ("entropy_visc", True, "type2"),
             ("fct", False, "type2"), ("fct", True, "type2")]
    outs, status = [], []
    for u, dt in zip(z["states"], z["dts"]):
        for (v, a, t) in modes:
            case.write_data(v, a, t)
            un, _, st = case.euler(u, dt)
            outs.append(un)
            status.append(st)
    runs = {}
    for tag, (v, a), n in (("viscous", ("viscous", False), 150), ("fct_alpha", ("fct", True), 100)):
        case.write_data(v, a, "type2")
        r = case.run(n)
        runs["run_%s_final" % tag] = r["final"]
        runs["run_%s_dts" % tag] = r["dts"]
        runs["run_%s_time" % tag] = r["time"]
    save("noround.npz", modes=np.array(["%s|%d|%s" % m for m in modes]),
         unext=np.stack(outs).reshape(2, len(modes), -1, 3), status=np.array(status).reshape(2, len(modes)), **runs)
    case.close()


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/ref_driver is not built: run `make -C oracle ref` first")
    only = set(sys.argv[1:])
    gens = [("lambda", gen_lambda), ("fct", gen_fct), ("par41", gen_par41_modes), ("par81", gen_par81_run),
            ("dambreak", gen_dambreak), ("dambreak9", gen_dambreak9), ("sgn", gen_sgn), ("bumps", gen_bumps), ("cav025", gen_cav025),
            ("cdisk", gen_cdisk_coriolis), ("friction", gen_par_friction), ("errors", gen_errors),
            ("flow", gen_flow), ("noround", gen_noround)]
    for name, fn in gens:
        if not only or name in only:
            fn()
    manifest = {f: os.path.getsize(os.path.join(HERE, f)) for f in sorted(os.listdir(HERE)) if f.endswith(".npz")}
    path = os.path.join(HERE, "MANIFEST.json")
    old = json.load(open(path)) if os.path.exists(path) else {}
    doc = {"generator": "tests/golden/make_golden.py", "reference": "t-bltg/shallow-water-eric built with "
           "amdflang -O2 -ffp-contract=off by oracle/Makefile", "seed": SEED, "files": manifest}
    if "copied" in old:
        doc["copied"] = old["copied"]
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=1)


if __name__ == "__main__":
    main()
