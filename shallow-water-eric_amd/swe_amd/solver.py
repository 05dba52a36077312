"""ctypes binding of libswe_amd.so (include/swe_amd.h): the host-side mirror of the reference's
MODULE update / MODULE fct interface (SOURCES/update.f90:5-10, SOURCES/fct.f90:5).

States are numpy arrays of shape (np, 3), which is the memory layout of the reference's
un(3,np) (h, qx, qy per node).  CSR arrays given to :class:`Solver` are 0-based; the C-ABI gets
the reference's 1-based numbering.  Every failure raises :class:`SweError`; the library has no
CPU fallback, so a missing or unloadable library is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libswe_amd.so")

SWE_OK, SWE_ENEG_H, SWE_ELAMBDA, SWE_EBADOPT, SWE_EHIP, SWE_EINVAL, SWE_ENOMEM, SWE_EFULL = range(8)
INSTANCE_RUNNING, INSTANCE_DONE, INSTANCE_FAILED = 1, 2, 3
NUM_KERNELS = 7
STEP_LAUNCHES, STEP_PERSISTENT = 0, 1
KERNELS = ("k_compute_dt", "k_alpha", "k_stage_lo", "k_fct_lo", "k_fct_rminus", "k_fct_apply", "k_ev_update")
GALERKIN, VISCOUS, ENTROPY_VISC, FCT = 0, 1, 2, 3
BC_NONE, BC_WALLS, BC_INFLOW, BC_VALUES, BC_STAGES = 0, 1, 2, 3, 4
VISCOSITY = {"galerkin": GALERKIN, "viscous": VISCOUS, "entropy_visc": ENTROPY_VISC, "fct": FCT}


class SweError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("swe status %d: %s" % (status, message))
        self.status = status


class Params(C.Structure):
    _fields_ = [("gravity", C.c_double), ("htiny", C.c_double), ("cfl", C.c_double),
                ("max_water_h", C.c_double), ("mannings", C.c_double),
                ("type_test", C.c_int), ("viscosity_type", C.c_int), ("viscous_type", C.c_int),
                ("alpha_limit", C.c_int), ("roundoff_fix", C.c_int), ("lambda_sgn", C.c_double)]


class BatchCaps(C.Structure):
    _fields_ = [("max_nodes", C.c_longlong), ("max_entries", C.c_longlong), ("max_boundary", C.c_longlong),
                ("max_instances", C.c_int)]


class MeshDesc(C.Structure):
    _fields_ = [("np", C.c_int), ("nnz", C.c_int), ("ia", C.c_void_p), ("ja", C.c_void_p),
                ("cx", C.c_void_p), ("cy", C.c_void_p), ("lumped", C.c_void_p), ("bath", C.c_void_p),
                ("n_h_dir", C.c_int), ("n_ux_dir", C.c_int), ("n_uy_dir", C.c_int),
                ("h_dir", C.c_void_p), ("ux_dir", C.c_void_p), ("uy_dir", C.c_void_p)]


# every symbol the header declares: name -> (restype, argtypes)
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
SIGNATURES = {
    "swe_last_error": (C.c_char_p, []),
    "swe_assemble_nnz": (C.c_long, [_I, _I, _P]),
    "swe_assemble": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "swe_assemble_device": (_I, [_I, _I, _P, _P, C.POINTER(_P), C.POINTER(C.c_longlong)]),
    "swe_assembly_get": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "swe_assembly_destroy": (_I, [_P]),
    "swe_create": (_I, [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(_P)]),
    "swe_create_from_assembly": (_I, [_P, _P, _I, _P, _I, _P, _I, _P, C.POINTER(Params), C.POINTER(_P)]),
    "swe_destroy": (_I, [_P]),
    "swe_errors_create": (_I, [_P, _I, _P, _P, C.POINTER(_P)]),
    "swe_compute_errors": (_I, [_P, C.c_double, _P]),
    "swe_ns_l1_ratio": (_I, [_P, _P]),
    "swe_errors_destroy": (_I, [_P]),
    "swe_batch_errors_create": (_I, [_P, _I, _I, _P, _P, C.POINTER(_P)]),
    "swe_set_params": (_I, [_P, C.POINTER(Params)]),
    "swe_set_bc": (_I, [_P, _I, _D, _D, _D]),
    "swe_set_bc_values": (_I, [_P, _P, _P, _P]),
    "swe_compute_dij": (_I, [_P, _P, _P]),
    "swe_compute_dt": (_I, [_P, _P, C.POINTER(_D)]),
    "swe_euler": (_I, [_P, _P, _P, _D]),
    "swe_fct_positivity": (_I, [_P, _P, _P, _P]),
    "swe_fct_positivity_csr": (_I, [_I, _I, _P, _P, _P, _P, _D, _P]),
    "swe_fct_positivity_csr_noalloc": (_I, [_I, _I, _P, _P, _P, _P, _D, _P, _P]),
    "swe_set_state": (_I, [_P, _P]),
    "swe_get_state": (_I, [_P, _P]),
    "swe_set_time": (_I, [_P, _D]),
    "swe_get_time": (_I, [_P, C.POINTER(_D)]),
    "swe_run": (_I, [_P, _I, C.POINTER(_D), _P]),
    "swe_run_async": (_I, [_P, _I]),
    "swe_sync": (_I, [_P]),
    "swe_stream": (_P, [_P]),
    "swe_enable_kernel_timing": (_I, [_P, C.c_char_p]),
    "swe_kernel_timing": (_I, [_P, C.c_char_p, C.POINTER(_I), C.POINTER(_D)]),
    "swe_tune": (_I, [_P, _I]),
    "swe_get_schedule": (_I, [_P, _P, _P]),
    "swe_set_schedule": (_I, [_P, _P, _P]),
    "swe_get_schedule_kinds": (_I, [_P, _P]),
    "swe_set_schedule_kinds": (_I, [_P, _P]),
    "swe_use_graph": (_I, [_P, _I]),
    "swe_set_step_form": (_I, [_P, _I]),
    "swe_get_step_form": (_I, [_P, _P]),
    "swe_fct_generic_csr": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "swe_stream_bandwidth": (_I, [C.c_ulonglong, _I, C.POINTER(_D), C.POINTER(_D)]),
    "swe_step_begin": (_I, [_P, C.POINTER(_D)]),
    "swe_step_finish": (_I, [_P, _P]),
    "swe_batch_bytes": (C.c_ulonglong, [C.POINTER(BatchCaps)]),
    "swe_batch_create": (_I, [C.POINTER(BatchCaps), C.POINTER(_P)]),
    "swe_batch_destroy": (_I, [_P]),
    "swe_batch_add": (_I, [_P, C.POINTER(MeshDesc), C.POINTER(Params), _P, _D, _D, C.c_longlong, C.POINTER(_I)]),
    "swe_batch_add_assembled": (_I, [_P, _P, _P, _I, _P, _I, _P, _I, _P, C.POINTER(Params), _P, _D, _D,
                                     C.c_longlong, C.POINTER(_I)]),
    "swe_batch_set_bc": (_I, [_P, _I, _I, _D, _D, _D]),
    "swe_batch_set_bc_values": (_I, [_P, _I, _P, _P, _P]),
    "swe_batch_remove": (_I, [_P, _I]),
    "swe_batch_step": (_I, [_P, _I]),
    "swe_batch_sync": (_I, [_P]),
    "swe_batch_status": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_D), C.POINTER(C.c_longlong), C.POINTER(_I)]),
    "swe_batch_poll": (_I, [_P, _P, _I, C.POINTER(_I)]),
    "swe_batch_get_state": (_I, [_P, _I, _P]),
    "swe_batch_step_async": (_I, [_P, _I, C.POINTER(C.c_longlong)]),
    "swe_batch_finished": (_I, [_P, C.c_longlong, _P, _I, C.POINTER(_I)]),
    "swe_batch_result": (_I, [_P, _I, _P, C.POINTER(_I), C.POINTER(_D), C.POINTER(C.c_longlong), C.POINTER(_I)]),
    "swe_batch_free": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(_I)]),
    "swe_batch_tune": (_I, [_P, _I]),
    "swe_batch_use_graph": (_I, [_P, _I]),
    "swe_batch_set_step_form": (_I, [_P, _I]),
    "swe_batch_get_step_form": (_I, [_P, _P]),
    "swe_batch_stream": (_P, [_P]),
    "swe_batch_enable_kernel_timing": (_I, [_P, C.c_char_p]),
    "swe_batch_kernel_timing": (_I, [_P, C.c_char_p, C.POINTER(_I), C.POINTER(_D)]),
}

_lib = None


def load_library(path: str | None = None):
    """Load libswe_amd.so and bind every declared symbol (raises if the library is missing).
    SWE_AMD_LIB names another build of the library to load instead of the in-tree one."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("SWE_AMD_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise SweError(SWE_EHIP, "libswe_amd.so is not built at %s (run __graft_entry__.build())" % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != SWE_OK:
        msg = load_library().swe_last_error()
        raise SweError(rc, msg.decode() if msg else "")


def assemble(jj0: np.ndarray, rr: np.ndarray, device: bool = False):
    """P1 set-up (construct_matrices): returns 0-based (ia, ja, diag), mass, lumped, cij (nnz,2).
    device=True assembles on the GPU (swe_assemble_device; the same bits)."""
    lib = load_library()
    jj1 = np.ascontiguousarray(jj0, dtype=np.int32) + 1
    rr = np.ascontiguousarray(rr, dtype=np.float64)
    npn, me = rr.shape[0], jj1.shape[0]
    if device:
        h = C.c_void_p()
        n = C.c_longlong()
        _check(lib.swe_assemble_device(npn, me, _ptr(jj1), _ptr(rr), C.byref(h), C.byref(n)))
        try:
            nnz = n.value
            ia = np.zeros(npn + 1, np.int32)
            ja = np.zeros(nnz, np.int32)
            diag = np.zeros(npn, np.int32)
            mass = np.zeros(nnz)
            lumped = np.zeros(npn)
            cx = np.zeros(nnz)
            cy = np.zeros(nnz)
            _check(lib.swe_assembly_get(h, _ptr(ia), _ptr(ja), _ptr(diag), _ptr(mass), _ptr(lumped), _ptr(cx),
                                        _ptr(cy)))
        finally:
            lib.swe_assembly_destroy(h)
        return ia - 1, ja - 1, diag - 1, mass, lumped, np.stack([cx, cy], 1)
    nnz = lib.swe_assemble_nnz(npn, me, _ptr(jj1))
    if nnz < 0:
        raise SweError(SWE_EINVAL, "invalid element list")
    ia = np.zeros(npn + 1, np.int32)
    ja = np.zeros(nnz, np.int32)
    diag = np.zeros(npn, np.int32)
    mass = np.zeros(nnz)
    lumped = np.zeros(npn)
    cx = np.zeros(nnz)
    cy = np.zeros(nnz)
    _check(lib.swe_assemble(npn, me, _ptr(jj1), _ptr(rr), _ptr(ia), _ptr(ja), _ptr(diag), _ptr(mass),
                            _ptr(lumped), _ptr(cx), _ptr(cy)))
    return ia - 1, ja - 1, diag - 1, mass, lumped, np.stack([cx, cy], 1)


def fct_positivity_csr(ia, ja, unext_scaled, f, htiny):
    """FCT_positivity (fct.f90:5-42) on a 0-based CSR matrix, without a solver handle."""
    lib = load_library()
    ia1 = np.ascontiguousarray(ia, np.int32) + 1
    ja1 = np.ascontiguousarray(ja, np.int32) + 1
    us = np.ascontiguousarray(unext_scaled, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    lij = np.zeros(ja1.size)
    _check(lib.swe_fct_positivity_csr(ia1.size - 1, ja1.size, _ptr(ia1), _ptr(ja1), _ptr(us), _ptr(f),
                                      float(htiny), _ptr(lij)))
    return lij


def fct_generic_csr(ia, ja, unext, maxn, minn, f, dg):
    """FCT_generic (SOURCES/SAFE/fct.f90:3-43) on the GPU; 0-based CSR (ia, ja); returns the limited
    unext (a new array)."""
    lib = load_library()
    ia1 = np.ascontiguousarray(ia, np.int32) + 1
    ja1 = np.ascontiguousarray(ja, np.int32) + 1
    u = np.array(unext, dtype=np.float64, copy=True)
    args = [np.ascontiguousarray(x, np.float64) for x in (maxn, minn, f, dg)]
    _check(lib.swe_fct_generic_csr(ia1.size - 1, ja1.size, _ptr(ia1), _ptr(ja1), _ptr(u), *[_ptr(a) for a in args]))
    return u


def make_params(*, gravity, htiny, cfl, max_water_h, viscosity="viscous", viscous_type=2, alpha_limit=False,
                type_test=5, mannings=0.0, roundoff_fix=True, lambda_sgn=0.0) -> Params:
    vt = VISCOSITY[viscosity] if isinstance(viscosity, str) else int(viscosity)
    return Params(gravity, htiny, cfl, max_water_h, mannings, type_test, vt, viscous_type,
                  int(bool(alpha_limit)), int(bool(roundoff_fix)), float(lambda_sgn))


def syst_size(type_test):
    """Components of the state (inputs%syst_size, SOURCES/main.f90:22-26): 5 for the SGN model of
    test 13 (h, qx, qy, h*eta, h*w), 3 otherwise."""
    return 5 if int(type_test) == 13 else 3


class MeshArrays:
    """A mesh in the C-ABI's form (1-based CSR, c_ij split, Dirichlet lists), kept alive with the
    descriptor that points into it."""

    def __init__(self, ia, ja, cij, lumped, bath, h_dir=(), ux_dir=(), uy_dir=()):
        self.ia1 = np.ascontiguousarray(ia, np.int32) + 1
        self.ja1 = np.ascontiguousarray(ja, np.int32) + 1
        self.np = self.ia1.size - 1
        self.nnz = self.ja1.size
        cij = np.asarray(cij, np.float64).reshape(self.nnz, 2)
        self.cx = np.ascontiguousarray(cij[:, 0])
        self.cy = np.ascontiguousarray(cij[:, 1])
        self.lumped = np.ascontiguousarray(lumped, np.float64)
        self.bath = np.ascontiguousarray(bath, np.float64)
        self.lists = [np.ascontiguousarray(x, np.int32) + 1 for x in (h_dir, ux_dir, uy_dir)]
        self.desc = MeshDesc(self.np, self.nnz, _ptr(self.ia1), _ptr(self.ja1), _ptr(self.cx), _ptr(self.cy),
                             _ptr(self.lumped), _ptr(self.bath),
                             self.lists[0].size, self.lists[1].size, self.lists[2].size,
                             _ptr(self.lists[0]), _ptr(self.lists[1]), _ptr(self.lists[2]))


class Solver:
    """One mesh, one parameter set, one device-resident state.

    Methods follow the reference: :meth:`compute_dij` (update.f90:291), :meth:`compute_dt`
    (COMPUTE_DT, main.f90:166), :meth:`euler` (update.f90:134), :meth:`fct_positivity`
    (fct.f90:5) and the SSP-RK3 time loop :meth:`run` (main.f90:61-92).
    """

    def __init__(self, ia, ja, cij, lumped, bath, *, gravity, htiny, cfl, max_water_h,
                 viscosity="viscous", viscous_type=2, alpha_limit=False, type_test=5, mannings=0.0,
                 roundoff_fix=True, lambda_sgn=0.0, h_dir=(), ux_dir=(), uy_dir=()):
        lib = load_library()
        self.mesh = MeshArrays(ia, ja, cij, lumped, bath, h_dir, ux_dir, uy_dir)
        self.np = self.mesh.np
        self.nnz = self.mesh.nnz
        self.nc = syst_size(type_test)
        self.params = make_params(gravity=gravity, htiny=htiny, cfl=cfl, max_water_h=max_water_h,
                                  viscosity=viscosity, viscous_type=viscous_type, alpha_limit=alpha_limit,
                                  type_test=type_test, mannings=mannings, roundoff_fix=roundoff_fix,
                                  lambda_sgn=lambda_sgn)
        h = C.c_void_p()
        _check(lib.swe_create(C.byref(self.mesh.desc), C.byref(self.params), C.byref(h)))
        self.h = h

    @classmethod
    def from_elements(cls, jj, rr, bath, *, h_dir=(), ux_dir=(), uy_dir=(), **params):
        """Set-up and solver on the GPU: construct_matrices by swe_assemble_device, then the solver
        built from the device assembly (swe_create_from_assembly); the matrices stay on the device.
        jj (me, 3) and the Dirichlet lists are 0-based; `params` are those of the constructor."""
        lib = load_library()
        jj1 = np.ascontiguousarray(jj, dtype=np.int32) + 1
        rr = np.ascontiguousarray(rr, dtype=np.float64)
        bath = np.ascontiguousarray(bath, dtype=np.float64)
        if bath.shape != (rr.shape[0],):
            raise SweError(SWE_EINVAL, "bath must have shape (%d,)" % rr.shape[0])
        lists = [np.ascontiguousarray(x, np.int32) + 1 for x in (h_dir, ux_dir, uy_dir)]
        self = cls.__new__(cls)
        self.h = None
        self.mesh = None
        self.params = make_params(**params)
        self.nc = syst_size(self.params.type_test)
        a = C.c_void_p()
        n = C.c_longlong()
        _check(lib.swe_assemble_device(rr.shape[0], jj1.shape[0], _ptr(jj1), _ptr(rr), C.byref(a), C.byref(n)))
        try:
            h = C.c_void_p()
            _check(lib.swe_create_from_assembly(a, _ptr(bath), lists[0].size, _ptr(lists[0]), lists[1].size,
                                                _ptr(lists[1]), lists[2].size, _ptr(lists[2]),
                                                C.byref(self.params), C.byref(h)))
        finally:
            lib.swe_assembly_destroy(a)
        self.h = h
        self.np = rr.shape[0]
        self.nnz = n.value
        return self

    # -- lifetime -------------------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            load_library().swe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- configuration ----------------------------------------------------------------------------
    def set_params(self, **kw):
        for k, v in kw.items():
            if k == "viscosity":
                k, v = "viscosity_type", VISCOSITY[v] if isinstance(v, str) else v
            setattr(self.params, k, v)
        _check(load_library().swe_set_params(self.h, C.byref(self.params)))

    def set_bc(self, kind, amp=0.0, period=1.0, level=0.0):
        _check(load_library().swe_set_bc(self.h, kind, amp, period, level))

    def set_bc_values(self, h_vals=None, ux_vals=None, uy_vals=None):
        arrs = [None if v is None else np.ascontiguousarray(v, np.float64) for v in (h_vals, ux_vals, uy_vals)]
        _check(load_library().swe_set_bc_values(self.h, *[_ptr(a) for a in arrs]))

    # -- reference routines on host arrays ----------------------------------------------------------
    def _state(self, un):
        un = np.ascontiguousarray(un, np.float64)
        if un.shape != (self.np, self.nc):
            raise SweError(SWE_EINVAL, "state must have shape (%d, %d), got %s" % (self.np, self.nc, un.shape))
        return un

    def compute_dij(self, un):
        un = self._state(un)
        d = np.zeros(self.nnz)
        _check(load_library().swe_compute_dij(self.h, _ptr(un), _ptr(d)))
        return d

    def compute_dt(self, un=None):
        dt = C.c_double()
        _check(load_library().swe_compute_dt(self.h, _ptr(None if un is None else self._state(un)), C.byref(dt)))
        return dt.value

    def euler(self, un, dt):
        un = self._state(un)
        out = np.zeros_like(un)
        _check(load_library().swe_euler(self.h, _ptr(un), _ptr(out), float(dt)))
        return out

    def fct_positivity(self, unext_scaled, f):
        us = np.ascontiguousarray(unext_scaled, np.float64)
        f = np.ascontiguousarray(f, np.float64)
        if us.size != self.np or f.size != self.nnz:
            raise SweError(SWE_EINVAL, "fct_positivity: sizes do not match the mesh")
        lij = np.zeros(self.nnz)
        _check(load_library().swe_fct_positivity(self.h, _ptr(us), _ptr(f), _ptr(lij)))
        return lij

    # -- device-resident time loop ----------------------------------------------------------------
    def set_state(self, un):
        _check(load_library().swe_set_state(self.h, _ptr(self._state(un))))

    def get_state(self):
        out = np.zeros((self.np, self.nc))
        _check(load_library().swe_get_state(self.h, _ptr(out)))
        return out

    @property
    def time(self):
        t = C.c_double()
        _check(load_library().swe_get_time(self.h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        _check(load_library().swe_set_time(self.h, float(t)))

    def run(self, nsteps, time=None):
        """nsteps SSP-RK3 steps; returns (end time, dt of each step)."""
        dts = np.zeros(max(int(nsteps), 1))
        t = C.c_double(self.time if time is None else float(time))
        _check(load_library().swe_run(self.h, int(nsteps), C.byref(t), _ptr(dts)))
        return t.value, dts[:nsteps]

    def step_begin(self):
        """COMPUTE_DT on the device-resident state; returns dt of the step it begins."""
        dt = C.c_double()
        _check(load_library().swe_step_begin(self.h, C.byref(dt)))
        return dt.value

    def step_finish(self, vals=None):
        """The three stages of the begun step; vals: (3, nh+nux+nuy) Dirichlet values per stage for
        the BC_STAGES program (h list, ux list, uy list)."""
        v = None if vals is None else np.ascontiguousarray(vals, np.float64)
        _check(load_library().swe_step_finish(self.h, _ptr(v)))

    def run_async(self, nsteps):
        _check(load_library().swe_run_async(self.h, int(nsteps)))

    def sync(self):
        _check(load_library().swe_sync(self.h))

    @property
    def stream(self):
        return load_library().swe_stream(self.h)

    def enable_kernel_timing(self, kernel):
        _check(load_library().swe_enable_kernel_timing(self.h, None if kernel is None else kernel.encode()))

    def kernel_timing(self, kernel):
        n = C.c_int()
        ms = C.c_double()
        _check(load_library().swe_kernel_timing(self.h, kernel.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    # -- launch schedules ---------------------------------------------------------------------------
    def tune(self, steps=3):
        """Autotune the launch schedule of every kernel on this mesh and state (state restored)."""
        _check(load_library().swe_tune(self.h, int(steps)))
        return self.schedule

    @property
    def schedule(self):
        """{kernel: (workgroup size, XCD-aware order, schedule kind)}."""
        b = np.zeros(NUM_KERNELS, np.int32)
        r = np.zeros(NUM_KERNELS, np.int32)
        kd = np.zeros(NUM_KERNELS, np.int32)
        _check(load_library().swe_get_schedule(self.h, _ptr(b), _ptr(r)))
        _check(load_library().swe_get_schedule_kinds(self.h, _ptr(kd)))
        return {k: (int(b[i]), int(r[i]), int(kd[i])) for i, k in enumerate(KERNELS)}

    @schedule.setter
    def schedule(self, sched):
        """Takes (workgroup size, XCD-aware order) or (workgroup size, XCD-aware order, kind) per
        kernel; a missing kind is 0."""
        b = np.array([sched[k][0] for k in KERNELS], np.int32)
        r = np.array([sched[k][1] for k in KERNELS], np.int32)
        kd = np.array([sched[k][2] if len(sched[k]) > 2 else 0 for k in KERNELS], np.int32)
        _check(load_library().swe_set_schedule(self.h, _ptr(b), _ptr(r)))
        _check(load_library().swe_set_schedule_kinds(self.h, _ptr(kd)))

    def use_graph(self, on=True):
        _check(load_library().swe_use_graph(self.h, int(bool(on))))

    @property
    def step_form(self):
        """STEP_LAUNCHES (kernel launches, captured graph per step) or STEP_PERSISTENT (fct: one
        cooperative kernel per run of steps, grid barriers between the passes)."""
        f = C.c_int()
        _check(load_library().swe_get_step_form(self.h, C.byref(f)))
        return f.value

    @step_form.setter
    def step_form(self, form):
        _check(load_library().swe_set_step_form(self.h, int(form)))

    def error_evaluator(self, jj, rr):
        """compute_errors (main.f90:198-263) on the device state; jj (me, 3) 0-based, rr (np, 2)."""
        return ErrorEvaluator(self, jj, rr)


class ErrorEvaluator:
    """The quadrature norms of compute_errors for one solver (swe_errors_create) or for instance `k`
    of a batch (swe_batch_errors_create)."""

    def __init__(self, owner, jj, rr, k=None):
        jj1 = np.ascontiguousarray(jj, np.int32) + 1
        rr = np.ascontiguousarray(rr, np.float64)
        npn = owner.np if k is None else owner._np[k][0]
        if rr.shape != (npn, 2) or jj1.ndim != 2 or jj1.shape[1] != 3:
            raise SweError(SWE_EINVAL, "jj must be (me, 3) and rr (%d, 2)" % npn)
        self.owner = owner                      # the handle must outlive the evaluator
        self.k = k
        e = C.c_void_p()
        if k is None:
            _check(load_library().swe_errors_create(owner.h, jj1.shape[0], _ptr(jj1), _ptr(rr), C.byref(e)))
        else:
            _check(load_library().swe_batch_errors_create(owner.h, int(k), jj1.shape[0], _ptr(jj1), _ptr(rr),
                                                          C.byref(e)))
        self.e = e

    def sums(self, time=None):
        """(3, 4): per component (h, qx, qy) the L1 error, L1 norm, L2 error, L2 norm at `time`
        (default: the solver's or the instance's time)."""
        out = np.zeros(12)
        if time is not None:
            t = float(time)
        elif self.k is None:
            t = self.owner.time
        else:
            t = self.owner.status(self.k)[1]
        _check(load_library().swe_compute_errors(self.e, t, _ptr(out)))
        return out.reshape(3, 4)

    def ns_l1_ratio(self):
        """The time loop's unit-10 diagnostic of tests 1, 2, 10 (main.f90:94-99): [q1, q2, q3, ratio],
        the L1 norms of qx, qy and h*sqrt(h) and 0.5*(q1+q2)/(g*q3)."""
        out = np.zeros(4)
        _check(load_library().swe_ns_l1_ratio(self.e, _ptr(out)))
        return out

    def relative(self, time=None):
        """What compute_errors prints: the relative L1 and L2 errors on h (main.f90:214,221), and
        the relative errors on q of its commented-out lines ((err_x+err_y)/(norm_x+norm_y))."""
        s = self.sums(time)
        return {"l1_h": s[0, 0] / s[0, 1], "l2_h": s[0, 2] / s[0, 3],
                "l1_q": (s[1, 0] + s[2, 0]) / (s[1, 1] + s[2, 1]), "l2_q": (s[1, 2] + s[2, 2]) / (s[1, 3] + s[2, 3])}

    def close(self):
        if getattr(self, "e", None):
            load_library().swe_errors_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
