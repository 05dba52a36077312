"""Batches: many meshes resident in one device arena and advanced together (swe_batch_* of
include/swe_amd.h).

Each instance runs the reference's time loop (SOURCES/main.f90:61-92) with its own parameters,
boundary program, dt, time and stopping rule; the results of an instance are bitwise those of
the same mesh run alone with :class:`swe_amd.solver.Solver`.  Instances share the scheme
(viscosity type, viscous type, alpha limiter, roundoff fix).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .solver import (BC_NONE, INSTANCE_DONE, INSTANCE_FAILED, INSTANCE_RUNNING, KERNELS, SWE_EFULL, BatchCaps,
                     MeshArrays, SweError, _check, _ptr, load_library, make_params, syst_size)

__all__ = ["Batch", "INSTANCE_RUNNING", "INSTANCE_DONE", "INSTANCE_FAILED"]


class Batch:
    """A device arena of meshes.  Capacities: node rows, entry slots, Dirichlet entries and
    instance records; :meth:`for_budget` derives them from a byte budget."""

    def __init__(self, max_nodes, max_entries, max_boundary, max_instances):
        lib = load_library()
        self.caps = BatchCaps(int(max_nodes), int(max_entries), int(max_boundary), int(max_instances))
        h = C.c_void_p()
        _check(lib.swe_batch_create(C.byref(self.caps), C.byref(h)))
        self.h = h
        self._np = {}

    @staticmethod
    def bytes_for(max_nodes, max_entries, max_boundary, max_instances):
        caps = BatchCaps(int(max_nodes), int(max_entries), int(max_boundary), int(max_instances))
        return int(load_library().swe_batch_bytes(C.byref(caps)))

    @classmethod
    def for_budget(cls, budget_bytes, max_instances=1024, entries_per_node=9.0, boundary_per_node=0.02):
        """The largest batch whose device arrays fit in budget_bytes, for meshes with about
        entries_per_node stored entries per node."""
        per_node = cls.bytes_for(1 << 20, int(entries_per_node * (1 << 20)), int(boundary_per_node * (1 << 20)),
                                 0) / float(1 << 20)
        fixed = cls.bytes_for(64, 1, 1, max_instances)
        nodes = max(64, int((budget_bytes - fixed) / per_node) // 64 * 64)
        return cls(nodes, int(entries_per_node * nodes), max(1, int(boundary_per_node * nodes)), max_instances)

    # -- lifetime ----------------------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            load_library().swe_batch_destroy(self.h)
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

    # -- instances -----------------------------------------------------------------------------------
    def add(self, mesh: MeshArrays, params, u0, *, t0=0.0, tfinal=math.inf, max_steps=2 ** 62,
            bc=(BC_NONE, 0.0, 1.0, 0.0)):
        """Admit a mesh; returns its id, or None when the batch has no room now."""
        u0 = np.ascontiguousarray(u0, np.float64)
        nc = syst_size(params.type_test)
        if u0.shape != (mesh.np, nc):
            raise SweError(5, "state must have shape (%d, %d)" % (mesh.np, nc))
        k = C.c_int(-1)
        lib = load_library()
        rc = lib.swe_batch_add(self.h, C.byref(mesh.desc), C.byref(params), _ptr(u0), float(t0), float(tfinal),
                               int(max_steps), C.byref(k))
        if rc == SWE_EFULL:
            return None
        _check(rc)
        if bc[0] != BC_NONE:
            _check(lib.swe_batch_set_bc(self.h, k.value, int(bc[0]), float(bc[1]), float(bc[2]), float(bc[3])))
        self._np[k.value] = (mesh.np, nc)
        return k.value

    def add_elements(self, jj, rr, bath, params, u0, *, h_dir=(), ux_dir=(), uy_dir=(), t0=0.0, tfinal=math.inf,
                     max_steps=2 ** 62, bc=(BC_NONE, 0.0, 1.0, 0.0)):
        """Admit a mesh given by its element list (jj (me, 3) and the Dirichlet lists 0-based, rr (np, 2)):
        construct_matrices and the layout run on the GPU (swe_assemble_device, swe_batch_add_assembled),
        while queued steps go on.  Returns its id, or None when the batch has no room now."""
        lib = load_library()
        jj1 = np.ascontiguousarray(jj, np.int32) + 1
        rr = np.ascontiguousarray(rr, np.float64)
        npn = rr.shape[0]
        bath = np.ascontiguousarray(bath, np.float64)
        u0 = np.ascontiguousarray(u0, np.float64)
        nc = syst_size(params.type_test)
        if bath.shape != (npn,) or u0.shape != (npn, nc):
            raise SweError(5, "bath must have shape (%d,) and the state (%d, %d)" % (npn, npn, nc))
        lists = [np.ascontiguousarray(x, np.int32) + 1 for x in (h_dir, ux_dir, uy_dir)]
        a = C.c_void_p()
        n = C.c_longlong()
        _check(lib.swe_assemble_device(npn, jj1.shape[0], _ptr(jj1), _ptr(rr), C.byref(a), C.byref(n)))
        k = C.c_int(-1)
        try:
            rc = lib.swe_batch_add_assembled(self.h, a, _ptr(bath), lists[0].size, _ptr(lists[0]), lists[1].size,
                                             _ptr(lists[1]), lists[2].size, _ptr(lists[2]), C.byref(params), _ptr(u0),
                                             float(t0), float(tfinal), int(max_steps), C.byref(k))
        finally:
            lib.swe_assembly_destroy(a)
        if rc == SWE_EFULL:
            return None
        _check(rc)
        if bc[0] != BC_NONE:
            _check(lib.swe_batch_set_bc(self.h, k.value, int(bc[0]), float(bc[1]), float(bc[2]), float(bc[3])))
        self._np[k.value] = (npn, nc)
        return k.value

    def add_problem(self, problem, *, tfinal=math.inf, max_steps=None, t0=0.0, device_setup=False):
        """Admit a :class:`swe_amd.problems.Problem` with its boundary program; device_setup=True
        assembles and lays the mesh out on the GPU from its element list (add_elements)."""
        hd, uxd, uyd = problem.dirichlet()
        prm = make_params(**problem.params)
        steps = problem.nsteps if max_steps is None else max_steps
        if device_setup:
            return self.add_elements(problem.mesh.jj, problem.mesh.rr, problem.bath, prm, problem.u0, h_dir=hd,
                                     ux_dir=uxd, uy_dir=uyd, t0=t0, tfinal=tfinal, max_steps=steps, bc=problem.bc)
        mesh = MeshArrays(problem.ia, problem.ja, problem.cij, problem.lumped, problem.bath, hd, uxd, uyd)
        return self.add(mesh, prm, problem.u0, t0=t0, tfinal=tfinal, max_steps=steps, bc=problem.bc)

    def remove(self, k):
        _check(load_library().swe_batch_remove(self.h, int(k)))
        self._np.pop(k, None)

    def set_bc_values(self, k, h_vals=None, ux_vals=None, uy_vals=None):
        arrs = [None if v is None else np.ascontiguousarray(v, np.float64) for v in (h_vals, ux_vals, uy_vals)]
        _check(load_library().swe_batch_set_bc_values(self.h, int(k), *[_ptr(a) for a in arrs]))

    # -- stepping ----------------------------------------------------------------------------------
    def step(self, nsteps=1):
        _check(load_library().swe_batch_step(self.h, int(nsteps)))

    def sync(self):
        _check(load_library().swe_batch_sync(self.h))

    def status(self, k):
        """(state, time, steps, error) of instance k."""
        st, err = C.c_int(), C.c_int()
        t = C.c_double()
        n = C.c_longlong()
        _check(load_library().swe_batch_status(self.h, int(k), C.byref(st), C.byref(t), C.byref(n), C.byref(err)))
        return st.value, t.value, n.value, err.value

    def poll(self, cap=4096):
        """Ids of the instances that are done or failed."""
        ids = np.zeros(cap, np.int32)
        n = C.c_int()
        _check(load_library().swe_batch_poll(self.h, _ptr(ids), cap, C.byref(n)))
        return [int(x) for x in ids[:n.value]]

    def error_evaluator(self, k, jj, rr):
        """compute_errors (main.f90:198-263) on instance k's device state; jj (me, 3) 0-based, rr (np, 2)."""
        from .solver import ErrorEvaluator
        return ErrorEvaluator(self, jj, rr, k=k)

    def get_state(self, k):
        out = np.zeros(self._np[k])
        _check(load_library().swe_batch_get_state(self.h, int(k), _ptr(out)))
        return out

    # -- pipelined serving ----------------------------------------------------------------------------
    def step_async(self, nsteps=1):
        """Queue nsteps steps and a copy of the instance table; returns the copy's ticket."""
        t = C.c_longlong()
        _check(load_library().swe_batch_step_async(self.h, int(nsteps), C.byref(t)))
        return t.value

    def finished(self, ticket, cap=4096):
        """Ids of the instances that the table copy of `ticket` shows done or failed (waits for
        that copy only; they take no later launches)."""
        ids = np.zeros(cap, np.int32)
        n = C.c_int()
        _check(load_library().swe_batch_finished(self.h, int(ticket), _ptr(ids), cap, C.byref(n)))
        return [int(x) for x in ids[:n.value]]

    def result(self, k):
        """(state, (np, 3) final state, time, steps, error) of an instance reported by
        :meth:`finished`, read while later steps run."""
        out = np.zeros(self._np[k])
        st, err = C.c_int(), C.c_int()
        t = C.c_double()
        n = C.c_longlong()
        _check(load_library().swe_batch_result(self.h, int(k), _ptr(out), C.byref(st), C.byref(t), C.byref(n),
                                               C.byref(err)))
        return st.value, out, t.value, n.value, err.value

    def free(self):
        """(largest free run of node rows, of entry slots, free instance records)."""
        r, e = C.c_longlong(), C.c_longlong()
        n = C.c_int()
        _check(load_library().swe_batch_free(self.h, C.byref(r), C.byref(e), C.byref(n)))
        return r.value, e.value, n.value

    @property
    def ids(self):
        return sorted(self._np)

    # -- schedules and timing ------------------------------------------------------------------------
    def tune(self, steps=3):
        _check(load_library().swe_batch_tune(self.h, int(steps)))

    def use_graph(self, on=True):
        _check(load_library().swe_batch_use_graph(self.h, int(bool(on))))

    @property
    def step_form(self):
        """As :attr:`swe_amd.solver.Solver.step_form`."""
        f = C.c_int()
        _check(load_library().swe_batch_get_step_form(self.h, C.byref(f)))
        return f.value

    @step_form.setter
    def step_form(self, form):
        _check(load_library().swe_batch_set_step_form(self.h, int(form)))

    @property
    def stream(self):
        return load_library().swe_batch_stream(self.h)

    def enable_kernel_timing(self, kernel):
        assert kernel is None or kernel in KERNELS
        _check(load_library().swe_batch_enable_kernel_timing(self.h, None if kernel is None else kernel.encode()))

    def kernel_timing(self, kernel):
        n = C.c_int()
        ms = C.c_double()
        _check(load_library().swe_batch_kernel_timing(self.h, kernel.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value
