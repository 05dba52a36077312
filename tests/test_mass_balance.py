"""Mass balance of the reference step on closed domains, measured on the oracle.

The total water volume is sum_i m_i h_i (lumped masses m_i).  On a domain closed by walls (qx = 0 on
the x sides, qy = 0 on the y sides, imposed by bdy after every stage) the Galerkin part of the mass
equation, sum_j c_ij . (v_j h_j) (smb_2_roundoff, update.f90:583-592), sums to zero over i, since
sum_i c_ij is the boundary integral of phi_j n and the normal discharge vanishes there.  What is left
is the graph viscosity, sum_i sum_j d_ij (H*_ji - H*_ij) with H*_ij = max(0, h_i + z_i - max(z_i, z_j))
(update.f90:568-591).  d_ij is built row by row, |c_ij| lambda(U_i, U_j, c_ij/|c_ij|) (update.f90:302-315),
and c_ji = -c_ij only on interior edges: on an edge between two wall nodes c_ij + c_ji is the boundary
integral of phi_i phi_j n, so d_ij != d_ji there and the first-order step does not conserve the
volume.  The fct scheme replaces d_ij by min(d_ij, r_ij) with the symmetric residual viscosity
r_ij = max(res_i, res_j) (update.f90:231-238, 870-876) wherever its limiter lets the high-order
fluxes through, so where the residual is the smaller (a wet domain) the step conserves the volume to
round-off.  The alpha limiter scales d_ij by max(alpha_i, alpha_j) (update.f90:784-787), which makes
d_ij the smaller one in smooth regions: the asymmetric wall viscosity comes back in, a drift of a
few 1e-12 relative over 200 steps.

What is checked here, on seeded 6,000-node Poisson-disk Delaunay meshes:
* one forward-Euler stage of the first-order scheme from rest changes the volume by exactly the
  wall-edge budget dt sum d_ij (H*_ji - H*_ij) (to the round-off of summing the volume), the interior
  edges cancel (lambda(U_i, U_j, n) = lambda(U_j, U_i, -n) and c_ji = -c_ij), and the change is far
  above round-off;
* 200 fct steps (no alpha limiter) on a wet basin and on a dam break over bumps keep the volume to
  1e-14 relative, while the first-order scheme on the same domains drifts by more than 1e-10.

The GPU path is held to the oracle's drift step by step in test_gpu_mass.py.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from swe_amd import problems


def _oracle(p, viscosity, alpha):
    pr = p.params
    return O.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=pr["gravity"], htiny=pr["htiny"], cfl=pr["cfl"],
                    max_water_h=pr["max_water_h"], viscosity_type=viscosity, viscous_type=2, alpha_limit=alpha,
                    type_test=pr["type_test"], mannings=pr["mannings"])


def volume(lumped, h):
    return math.fsum(lumped * h)


def wall_nodes(p):
    _, uxd, uyd = p.dirichlet()
    w = np.zeros(p.np, bool)
    w[np.concatenate([uxd, uyd])] = True
    return w


@pytest.fixture(scope="module")
def basin():
    return problems.family_member(6000, 3, kind="basin")


@pytest.fixture(scope="module")
def dam():
    return problems.family_member(6000, 5, kind="dam")


def test_first_order_volume_change_is_the_wall_edge_budget(oracle, basin):
    p = basin
    u = p.u0.copy()
    assert np.all(u[:, 0] > 0.25) and np.all(u[:, 1:] == 0.0)
    o = _oracle(p, O.VISCOUS, False)
    try:
        dt, err = o.compute_dt(u)
        assert err == 0
        out, d, err = o.euler(u, dt)
        assert err == 0
    finally:
        o.close()
    v0 = volume(p.lumped, u[:, 0])
    dv = volume(p.lumped, out[:, 0]) - v0
    rows = np.repeat(np.arange(p.np), np.diff(p.ia))
    h, z = u[:, 0], p.bath
    zm = np.maximum(z[rows], z[p.ja])
    hs_ij = np.maximum(0.0, h[rows] + z[rows] - zm)
    hs_ji = np.maximum(0.0, h[p.ja] + z[p.ja] - zm)
    term = d * (hs_ji - hs_ij)
    off = rows != p.ja
    w = wall_nodes(p)
    wall_edge = off & w[rows] & w[p.ja]
    wall = dt * math.fsum(term[wall_edge])
    interior = dt * math.fsum(term[off & ~wall_edge])
    eps = 1e-15 * v0
    print("volume %.6g, change %.6e, wall budget %.6e, interior %.3e" % (v0, dv, wall, interior))
    assert abs(interior) <= eps
    assert abs(dv - wall) <= eps
    assert abs(dv) >= 100 * eps


@pytest.mark.parametrize("case", ["basin", "dam"])
def test_fct_keeps_the_volume_of_a_closed_wet_domain(oracle, case, request):
    p = request.getfixturevalue(case)
    hd, uxd, uyd = p.dirichlet()
    kind, amp, period, level = p.bc
    assert hd.size == 0 and kind == problems.BC_WALLS
    v0 = volume(p.lumped, p.u0[:, 0])
    drift = {}
    for visc in (O.FCT, O.VISCOUS):
        o = _oracle(p, visc, False)
        try:
            u, t, worst = p.u0.copy(), 0.0, 0.0
            for _ in range(10):
                u, _, t, err = o.run(u, 20, t0=t, bc_kind=kind, amp=amp, period=period, level=level,
                                     nodes_ux=uxd, nodes_uy=uyd)
                assert err == 0
                worst = max(worst, abs(volume(p.lumped, u[:, 0]) - v0) / v0)
        finally:
            o.close()
        drift[visc] = worst
    print(case, "largest relative drift over 200 steps: fct %.3g, first order %.3g" % (drift[O.FCT], drift[O.VISCOUS]))
    assert drift[O.FCT] <= 1e-14
    assert drift[O.VISCOUS] >= 1e-10
