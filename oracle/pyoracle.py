"""ctypes wrapper of liboracle_swe.so and of the reference driver -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, bench.py's CPU-baseline leg and __graft_entry__.smoke() as the checker.
The product (shallow-water-eric_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_swe.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
# the same driver with entropy_residual's O(np^2) rescale loop reduced to its one assignment
# (oracle/hoist_rescale.awk): the same bits, timed as the CPU baseline
REF_DRIVER_HOIST = os.path.join(HERE, "_ref", "ref_driver_hoist")
# the reference with MODULE update's if_roundoff_fix = .FALSE. (the smb_2 branches of euler)
REF_DRIVER_NOROUND = os.path.join(HERE, "_ref", "ref_driver_noround")

GALERKIN, VISCOUS, ENTROPY_VISC, FCT = 0, 1, 2, 3
VISC_NAMES = {GALERKIN: "galerkin", VISCOUS: "viscous", ENTROPY_VISC: "entropy_visc", FCT: "fct"}
ORC_OK, ENEG_H, ELAMBDA, EBADOPT = 0, 1, 2, 3   # orc_run status (swe_oracle.h)


class OrcParams(C.Structure):
    _fields_ = [("gravity", C.c_double), ("htiny", C.c_double), ("cfl", C.c_double),
                ("max_water_h", C.c_double), ("mannings", C.c_double),
                ("type_test", C.c_int), ("viscosity_type", C.c_int), ("viscous_type", C.c_int),
                ("alpha_limit", C.c_int), ("roundoff_fix", C.c_int), ("lambda_sgn", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_swe.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liboracle_swe.so is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_csr_nnz.restype = C.c_long
        L.orc_csr_nnz.argtypes = [C.c_int, C.c_int, P]
        L.orc_assemble.restype = C.c_int
        L.orc_assemble.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, P, P, P, P, P, C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [P]
        L.orc_set_params.argtypes = [P, C.POINTER(OrcParams)]
        L.orc_syst_size.restype = C.c_int
        L.orc_syst_size.argtypes = [P]
        L.orc_set_coords.argtypes = [P, P]
        L.orc_lambda_batch.argtypes = [C.c_int, P, C.c_double, C.c_double, P, P]
        L.orc_fct_positivity.argtypes = [C.c_int, P, P, P, P, C.c_double, P]
        L.orc_fct_generic.argtypes = [C.c_int, P, P, P, P, P, P, P]
        L.orc_compute_dij.restype = C.c_int
        L.orc_compute_dij.argtypes = [P, P, P]
        L.orc_compute_dt.restype = C.c_double
        L.orc_compute_dt.argtypes = [P, P, P]
        L.orc_euler.restype = C.c_int
        L.orc_euler.argtypes = [P, P, P, C.c_double, P]
        L.orc_compute_errors.restype = C.c_int
        L.orc_compute_errors.argtypes = [C.c_int, C.c_int, P, P, P, C.c_int, C.c_double, C.c_double, C.c_double,
                                         P, P]
        L.orc_ns_l1_ratio.restype = None
        L.orc_ns_l1_ratio.argtypes = [C.c_int, C.c_int, P, P, P, C.c_int, C.c_double, P, P]
        L.orc_sol_anal.restype = C.c_double
        L.orc_sol_anal.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.POINTER(C.c_int)]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [P, P, C.c_int, P, P, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_int, P, C.c_int, P, C.c_int, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def assemble(jj0: np.ndarray, rr: np.ndarray):
    """P1 set-up of construct_matrices: returns (ia, ja, diag, mass, lumped, cij) 0-based."""
    jj0 = np.ascontiguousarray(jj0, dtype=np.int32)
    rr = np.ascontiguousarray(rr, dtype=np.float64)
    npn, me = rr.shape[0], jj0.shape[0]
    nnz = lib().orc_csr_nnz(npn, me, _p(jj0))
    ia = np.zeros(npn + 1, np.int32)
    ja = np.zeros(nnz, np.int32)
    diag = np.zeros(npn, np.int32)
    mass = np.zeros(nnz)
    lumped = np.zeros(npn)
    cij = np.zeros((nnz, 2))
    lib().orc_assemble(npn, me, _p(jj0), _p(rr), _p(ia), _p(ja), _p(diag), _p(mass), _p(lumped), _p(cij))
    return ia, ja, diag, mass, lumped, cij


def lambda_batch(rec8: np.ndarray, gravity: float, htiny: float):
    rec8 = np.ascontiguousarray(rec8, dtype=np.float64)
    out = np.zeros(rec8.shape[0])
    err = C.c_int(0)
    lib().orc_lambda_batch(rec8.shape[0], _p(rec8), gravity, htiny, _p(out), C.byref(err))
    return out, err.value


def fct_generic(ia, ja, f, unext, maxn, minn, dg):
    """FCT_generic (SAFE/fct.f90:3-43) on a 0-based CSR matrix; returns the limited unext."""
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    u = np.array(unext, dtype=np.float64, copy=True)
    args = [np.ascontiguousarray(x, np.float64) for x in (f, maxn, minn, dg)]
    lib().orc_fct_generic(ia.size - 1, _p(ia), _p(ja), _p(args[0]), _p(args[1]), _p(args[2]), _p(args[3]), _p(u))
    return u


def compute_errors(jj0, rr, u, type_test, t, gravity, htiny, terms=False):
    """compute_errors' quadrature norms (main.f90:198-263): array (3, 4) of [L1 error, L1 norm,
    L2 error, L2 norm] per component (h, qx, qy); with terms=True also every summand, (3, 4, 3*me)."""
    jj0 = np.ascontiguousarray(jj0, dtype=np.int32)
    rr = np.ascontiguousarray(rr, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    me = jj0.shape[0]
    out = np.zeros(12)
    tq = np.zeros(12 * 3 * me) if terms else None
    rc = lib().orc_compute_errors(rr.shape[0], me, _p(jj0), _p(rr), _p(u), int(type_test), float(t), float(gravity),
                                  float(htiny), _p(out), _p(tq) if terms else None)
    if rc != 0:
        raise ValueError("compute_errors: no analytic solution for test %d" % type_test)
    out = out.reshape(3, 4)
    return (out, tq.reshape(3, 4, 3 * me)) if terms else out


def ns_l1_ratio(jj0, rr, u, gravity, terms=False):
    """The time loop's unit-10 diagnostic (main.f90:94-99): [q1, q2, q3, 0.5*(q1+q2)/(g*q3)] with the
    ns_l1 norms of qx, qy and h*sqrt(h); u (np, nc) with h, qx, qy first; with terms=True also every
    summand, (3, 3*me)."""
    jj0 = np.ascontiguousarray(jj0, dtype=np.int32)
    rr = np.ascontiguousarray(rr, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    me = jj0.shape[0]
    out = np.zeros(4)
    tq = np.zeros(3 * 3 * me) if terms else None
    lib().orc_ns_l1_ratio(rr.shape[0], me, _p(jj0), _p(rr), _p(u), int(u.shape[1]), float(gravity), _p(out),
                          _p(tq) if terms else None)
    return (out, tq.reshape(3, 3 * me)) if terms else out


def sol_anal(type_test, k, x, y, t, gravity, htiny):
    ok = C.c_int(0)
    v = lib().orc_sol_anal(int(type_test), int(k), float(x), float(y), float(t), float(gravity), float(htiny),
                           C.byref(ok))
    if not ok.value:
        raise ValueError("sol_anal: test %d not restated" % type_test)
    return v


def fct_positivity(ia, ja, f, unext_scaled, htiny):
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    f = np.ascontiguousarray(f, np.float64)
    us = np.ascontiguousarray(unext_scaled, np.float64)
    lij = np.zeros(ja.size)
    lib().orc_fct_positivity(ia.size - 1, _p(ia), _p(ja), _p(f), _p(us), htiny, _p(lij))
    return lij


class Oracle:
    """A reference-step context over one mesh.  States are (np,3) arrays (rows = nodes)."""

    def __init__(self, ia, ja, cij, lumped, bath, *, gravity, htiny, cfl, max_water_h,
                 viscosity_type=VISCOUS, viscous_type=2, alpha_limit=False, type_test=5,
                 mannings=0.0, roundoff_fix=True, lambda_sgn=0.0, rr=None):
        """States are (np, nc): nc = 3, or 5 (h, qx, qy, h*eta, h*w) for the SGN model of
        type_test 13.  rr (np, 2): node coordinates, for the reference's bdy (bc_kind 3)."""
        self.ia = np.ascontiguousarray(ia, np.int32)
        self.ja = np.ascontiguousarray(ja, np.int32)
        self.np = self.ia.size - 1
        self.nnz = self.ja.size
        self.prm = OrcParams(gravity, htiny, cfl, max_water_h, mannings, type_test,
                             viscosity_type, viscous_type, int(alpha_limit), int(roundoff_fix), float(lambda_sgn))
        cij = np.ascontiguousarray(cij, np.float64).reshape(-1)
        self._keep = (self.ia, self.ja, cij, np.ascontiguousarray(lumped, np.float64),
                      np.ascontiguousarray(bath, np.float64))
        self.h = lib().orc_create(self.np, self.nnz, _p(self.ia), _p(self.ja), _p(self._keep[2]),
                                  _p(self._keep[3]), _p(self._keep[4]), C.byref(self.prm))
        self.nc = lib().orc_syst_size(self.h)
        if rr is not None:
            lib().orc_set_coords(self.h, _p(np.ascontiguousarray(rr, np.float64)))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute_dij(self, u):
        u = np.ascontiguousarray(u, np.float64)
        d = np.zeros(self.nnz)
        err = lib().orc_compute_dij(self.h, _p(u), _p(d))
        return d, err

    def compute_dt(self, u):
        u = np.ascontiguousarray(u, np.float64)
        err = C.c_int(0)
        dt = lib().orc_compute_dt(self.h, _p(u), C.byref(err))
        return dt, err.value

    def euler(self, u, dt):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        d = np.zeros(self.nnz)
        err = lib().orc_euler(self.h, _p(u), _p(out), dt, _p(d))
        return out, d, err

    def run(self, u, nsteps, t0=0.0, bc_kind=0, amp=0.0, period=1.0, level=0.0,
            nodes_h=(), nodes_ux=(), nodes_uy=()):
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        t = C.c_double(t0)
        dts = np.zeros(max(nsteps, 1))
        nh = np.ascontiguousarray(nodes_h, np.int32)
        nx = np.ascontiguousarray(nodes_ux, np.int32)
        ny = np.ascontiguousarray(nodes_uy, np.int32)
        err = lib().orc_run(self.h, _p(u), nsteps, C.byref(t), _p(dts), bc_kind, amp, period, level,
                            nh.size, _p(nh), nx.size, _p(nx), ny.size, _p(ny))
        return u, dts[:nsteps], t.value, err


# ---------------------------------------------------------------------------------------------
# The reference build (oracle/_ref/ref_driver): only where it was built, i.e. in the container
# that holds the reference sources.  Its outputs are committed as fixtures under tests/golden/.
# ---------------------------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def data_file(path, *, mesh_dir, mesh_file, test, tfinal=1.0, cfl=0.25, viscosity="viscous",
              viscous_type="type2", alpha=False, h_dir=(), ux_dir=(), uy_dir=(), mannings=None, lambda_sgn=None):
    """Write a reference `data` file (keys of SOURCES/read_data.f90:42-97)."""
    def lst(v):
        return " ".join(str(int(x)) for x in v) if len(v) else "1"
    lines = [
        "===Name of directory for mesh file===", "'%s'" % mesh_dir,
        "===Name of mesh file===", mesh_file,
        "===Number of subdomains in the mesh===", "1",
        "===List of subdomain in the mesh===", "1",
        "===Type of finite element===", "1",
        "===Final time===", repr(float(tfinal)),
        "===CFL number===", repr(float(cfl)),
        "===Viscosity type (galerkin, viscous, entropy_visc, fct)===", "'%s'" % viscosity,
        "===Type of first-order viscosity===", "'%s'" % viscous_type,
        "===Alpha limitation? (True/False)===", ".t." if alpha else ".f.",
        "===Lumping the mass matrix? (true/false)===", ".t.",
        "===Test case number===", str(int(test)),
        "===How many Dirichlet boundaries for h?===", str(len(h_dir)),
        "===List of Dirichlet boundaries for h?===", lst(h_dir),
        "===How many Dirichlet boundaries for ux?===", str(len(ux_dir)),
        "===List of Dirichlet boundaries for ux?===", lst(ux_dir),
        "===How many Dirichlet boundaries for uy?===", str(len(uy_dir)),
        "===List of Dirichlet boundaries for uy?===", lst(uy_dir),
    ]
    if mannings is not None:
        lines += ["===Mannings coefficient===", repr(float(mannings))]
    if lambda_sgn is not None:
        lines += ["===Lambda for SGN model===", repr(float(lambda_sgn))]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_custom_init(path, bath, u, *, gravity, max_water_h, bc_kind=0, amp=0.0, period=1.0, level=0.0):
    """Init file read by ref_driver custom_init: header then bath(np), un(3,np)."""
    npn = bath.size
    with open(path, "wb") as f:
        f.write(np.array([npn, bc_kind], dtype="<i4").tobytes())
        f.write(np.array([gravity, max_water_h, amp, period, level], dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(bath, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(u, dtype="<f8").tobytes())


def _big_stack():
    # the reference keeps (3, np) work arrays as automatic arrays on the stack: lift the soft stack
    # limit to the hard one in the child so that meshes beyond a few thousand nodes fit
    import resource
    _, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run_ref(args, cwd, timeout=3600, driver=None, pin=None):
    """Run the reference driver (or `driver`) in `cwd`; pin = CPU core to pin it to with taskset."""
    cmd = [driver or REF_DRIVER] + [str(a) for a in args]
    if pin is not None:
        cmd = ["taskset", "-c", str(pin)] + cmd
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=timeout, preexec_fn=_big_stack)
    if r.returncode != 0 or "BUG" in r.stdout or "STOP" in r.stdout:
        raise RuntimeError("ref_driver %s failed (%d):\n%s\n%s" % (args[0], r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    return r.stdout


def ref_case(d, problem, test, mesh_file="mesh.FEM"):
    """Write the mesh (.FEM), `data` file and initial state of a swe_amd.problems.Problem into the
    directory d for the reference driver; returns the init argument ("none" or the init file)."""
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "shallow-water-eric_amd"))
    from swe_amd import mesh as M
    pr = problem.params
    M.write_fem(os.path.join(d, mesh_file), problem.mesh)
    data_file(os.path.join(d, "data"), mesh_dir=".", mesh_file=mesh_file, test=test, cfl=pr["cfl"],
              viscosity=pr["viscosity"], viscous_type="type%d" % pr["viscous_type"], alpha=pr["alpha_limit"],
              h_dir=problem.h_sides, ux_dir=problem.ux_sides, uy_dir=problem.uy_sides,
              mannings=pr["mannings"] if test in (9, 10) else None,
              lambda_sgn=pr.get("lambda_sgn") if test == 13 else None)
    if problem.reference_init:
        return "none"
    kind, amp, period, level = problem.bc
    write_custom_init(os.path.join(d, "init.bin"), problem.bath, problem.u0, gravity=pr["gravity"],
                      max_water_h=pr["max_water_h"], bc_kind=kind, amp=amp, period=period, level=level)
    return "init.bin"


def ref_time_loop(d, init, nsteps, driver=None, pin=None, out="run", timeout=3600, nc=3):
    """nsteps of the driver's SSP-RK3 loop on the case in d: (final (np,nc), dts, end time, seconds of
    the time loop by the driver's own clock, set-up excluded)."""
    od = os.path.join(d, out)
    os.makedirs(od, exist_ok=True)
    run_ref(["run", "data", init, nsteps, out], cwd=d, timeout=timeout, driver=driver, pin=pin)
    final = np.fromfile(os.path.join(od, "final.f8"), dtype="<f8").reshape(-1, nc)
    return (final, np.fromfile(os.path.join(od, "dt.f8")), float(np.fromfile(os.path.join(od, "time.f8"))[0]),
            float(np.fromfile(os.path.join(od, "loop_seconds.f8"))[0]))


def read_setup(d):
    def i4(n):
        return np.fromfile(os.path.join(d, n), dtype="<i4")

    def f8(n):
        return np.fromfile(os.path.join(d, n), dtype="<f8")
    dims = i4("dims.i4")
    npn, me, nnz = int(dims[0]), int(dims[1]), int(dims[2])
    sc = f8("scalars.f8")
    nc = 5 if int(dims[6]) == 13 else 3
    return dict(np=npn, me=me, nnz=nnz, type_test=int(dims[6]), syst_size=nc,
                gravity=sc[0], htiny=sc[1], max_water_h=sc[2], cfl=sc[3], epsilon=sc[4], time0=sc[5],
                jj=i4("jj.i4").reshape(me, 3) - 1, rr=f8("rr.f8").reshape(npn, 2),
                ia=i4("ia.i4") - 1, ja=i4("ja.i4") - 1, mass=f8("mass.f8"), lumped=f8("lumped.f8"),
                diag=i4("diag.i4") - 1, cij=f8("cij.f8").reshape(nnz, 2), bath=f8("bath.f8"),
                u0=f8("u0.f8").reshape(npn, nc),
                h_js_D=i4("h_js_D.i4") - 1, ux_js_D=i4("ux_js_D.i4") - 1, uy_js_D=i4("uy_js_D.i4") - 1)


def tmpdir():
    return tempfile.mkdtemp(prefix="swe_", dir="/tmp")
