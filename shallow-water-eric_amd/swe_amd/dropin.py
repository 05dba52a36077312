"""The reference's time loop driving the per-stage drop-in entry points from the host.

This is what `SOURCES/main.f90` does when it is linked with the replacement MODULE update (the
Fortran binding in `fortran/`): each step calls `COMPUTE_DT` (`main.f90:166-174`: `compute_dij`
on the device, then `dt = CFL*1/MAXVAL(ABS(dij%aa(diag))/lumped)` on the host), then three
`euler` stages with the SSP-RK3 combinations and `bdy` on host arrays (`main.f90:71-91`), so the
state and `dij%aa` cross PCIe every stage.  It is the plain per-instance form the batched and
served forms are compared with, next to the device-resident `swe_run` loop, and its results are
bitwise those of the device-resident loop.
"""
from __future__ import annotations

import numpy as np

from .solver import BC_NONE, BC_WALLS, SweError

__all__ = ["DropinLoop"]


class DropinLoop:
    """Host time loop over one :class:`swe_amd.solver.Solver` built for `problem`.  Boundary
    programs: none, or walls (qx = 0 on the ux list, qy = 0 on the uy list, the reference's
    `bdy` with a zero `sol_anal` for the velocity components)."""

    def __init__(self, solver, problem):
        kind = problem.bc[0]
        if kind not in (BC_NONE, BC_WALLS):
            raise SweError(5, "the host loop handles no-boundary and wall programs only")
        self.s = solver
        self.diag = np.asarray(problem.diag, np.int64)
        self.lumped = np.asarray(problem.lumped, np.float64)
        self.cfl = float(problem.params["cfl"])
        _, ux, uy = problem.dirichlet()
        self.walls = kind == BC_WALLS
        self.ux = np.asarray(ux, np.int64)
        self.uy = np.asarray(uy, np.int64)

    def bdy(self, u):
        # bdy (main.f90:176-188) for the wall program
        if self.walls:
            u[self.ux, 1] = 0.0
            u[self.uy, 2] = 0.0
        return u

    def compute_dt(self, u):
        # COMPUTE_DT (main.f90:166-174)
        d = self.s.compute_dij(u)
        return self.cfl * 1 / np.max(np.abs(d[self.diag]) / self.lumped)

    def step(self, un, time):
        """One step of main.f90:71-91; returns (u^{n+1}, time + dt, dt)."""
        dt = self.compute_dt(un)
        uo = un
        u = self.bdy(self.s.euler(uo, dt))
        ui = self.s.euler(u, dt)
        u = self.bdy((3 * uo + ui) / 4)
        ui = self.s.euler(u, dt)
        u = self.bdy((uo + 2 * ui) / 3)
        return u, time + dt, dt

    def run(self, u0, nsteps, time=0.0):
        u = np.ascontiguousarray(u0, np.float64)
        dts = []
        for _ in range(int(nsteps)):
            u, time, dt = self.step(u, time)
            dts.append(dt)
        return u, time, np.array(dts)
