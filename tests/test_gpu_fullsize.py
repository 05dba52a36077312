"""The workloads at their full sizes: the 1,006,209-node dam break (FCT + alpha + Manning, walls), the
16,785,409-node bumps problem on the cell-split jittered lattice and the same problem on the
16.7M-node Poisson-disk Delaunay mesh (FCT + alpha, sinusoidal inflow, dry islands).

The oracle is too slow for whole runs at these sizes, so: the first steps against the oracle on the
full mesh (bitwise for the bumps case; 1e-11 relative with Manning friction, whose pow differs from
glibc's in rare last bits), then the whole run through size-independent properties -- depth >= 0 and
finite everywhere, no error flag, and every bit of the final state and of every dt the same for a solver
laid out on the host (reverse Cuthill-McKee numbering) and one assembled and laid out on the device
(the caller's numbering), i.e. the result does not depend on the processing order."""
import numpy as np
import pytest

from conftest import rel_diff, same_bits

pytestmark = pytest.mark.gpu

TOL = 1e-11


@pytest.fixture(scope="module")
def dam_break():
    from swe_amd import problems
    p = problems.dam_break()
    print("dam break", p.np, flush=True)
    return p


@pytest.fixture(scope="module")
def bumps():
    from swe_amd import problems
    p = problems.bumps()
    print("bumps", p.np, flush=True)
    return p


def _oracle_for(oracle, p):
    pr = p.params
    return oracle.Oracle(p.ia, p.ja, p.cij, p.lumped, p.bath, gravity=pr["gravity"], htiny=pr["htiny"],
                         cfl=pr["cfl"], max_water_h=pr["max_water_h"], viscosity_type=oracle.FCT,
                         alpha_limit=pr["alpha_limit"], type_test=pr["type_test"], mannings=pr["mannings"])


def _oracle_run(oracle, p, nsteps):
    hd, uxd, uyd = p.dirichlet()
    kind, amp, period, level = p.bc
    orc = _oracle_for(oracle, p)
    try:
        return orc.run(p.u0, nsteps, bc_kind=kind, amp=amp, period=period, level=level, nodes_h=hd,
                       nodes_ux=uxd, nodes_uy=uyd)
    finally:
        orc.close()


def _run(p, nsteps, device_setup=False):
    with p.solver(device_setup=device_setup) as s:
        s.set_state(p.u0)
        t, dts = s.run(nsteps, time=0.0)
        return s.get_state(), t, dts


def test_dam_break_full_size_first_steps_against_oracle(oracle, dam_break):
    p = dam_break
    ref, rdts, rt, err = _oracle_run(oracle, p, 10)
    assert err == 0
    u, t, dts = _run(p, 10)
    l1, linf = rel_diff(u, ref)
    print("dam break 1M, 10 steps: rel L1 %.3g Linf %.3g, bitwise fraction %.6f" % (l1, linf, np.mean(u == ref)),
          flush=True)
    assert l1 <= TOL and linf <= TOL
    assert np.abs(dts - rdts).max() <= TOL * rdts.max() and abs(t - rt) <= TOL * rt


def test_dam_break_full_size_500_steps_properties(dam_break):
    p = dam_break
    u, t, dts = _run(p, 500)
    print("dam break 1M, 500 steps: t = %.6g, min h %.3g" % (t, u[:, 0].min()), flush=True)
    assert np.all(np.isfinite(u)) and u[:, 0].min() >= 0.0 and np.all(dts > 0)
    u2, t2, dts2 = _run(p, 500, device_setup=True)
    assert same_bits(u, u2) and same_bits(dts, dts2) and t == t2


def test_bumps_16m_first_step_against_oracle(oracle, bumps):
    p = bumps
    ref, rdts, rt, err = _oracle_run(oracle, p, 1)
    assert err == 0
    u, t, dts = _run(p, 1)
    print("bumps 16.8M, 1 step: bitwise fraction %.6f" % np.mean(u == ref), flush=True)
    assert same_bits(dts, rdts) and t == rt
    assert same_bits(u, ref)


def test_bumps_16m_200_steps_properties(bumps):
    p = bumps
    u, t, dts = _run(p, 200)
    print("bumps 16.8M, 200 steps: t = %.6g, min h %.3g, dry nodes %d" % (t, u[:, 0].min(), int(np.sum(u[:, 0] == 0))),
          flush=True)
    assert np.all(np.isfinite(u)) and u[:, 0].min() >= 0.0 and np.all(dts > 0)
    u2, t2, dts2 = _run(p, 200, device_setup=True)
    assert same_bits(u, u2) and same_bits(dts, dts2) and t == t2


@pytest.fixture(scope="module")
def bumps_unstructured():
    from swe_amd import problems
    p = problems.bumps_unstructured()
    print("bumps unstructured", p.np, p.nnz, flush=True)
    return p


def test_bumps_unstructured_16m_first_step_against_oracle(oracle, bumps_unstructured):
    """The 16.7M-node Poisson-disk Delaunay mesh (the headline workload): one step on the whole mesh,
    every bit and the dt against the oracle, for the device set-up (Hilbert locality order)."""
    p = bumps_unstructured
    ref, rdts, rt, err = _oracle_run(oracle, p, 1)
    assert err == 0
    u, t, dts = _run(p, 1, device_setup=True)
    print("bumps unstructured 16.7M, 1 step: bitwise fraction %.6f" % np.mean(u == ref), flush=True)
    assert same_bits(dts, rdts) and t == rt
    assert same_bits(u, ref)


def test_bumps_unstructured_16m_200_steps_properties(bumps_unstructured):
    """The whole 200-step run: depth >= 0 and finite, and the same bits for the host layout (reverse
    Cuthill-McKee order) and the device layout (Hilbert order)."""
    p = bumps_unstructured
    u, t, dts = _run(p, 200, device_setup=True)
    print("bumps unstructured 16.7M, 200 steps: t = %.6g, min h %.3g, dry nodes %d"
          % (t, u[:, 0].min(), int(np.sum(u[:, 0] == 0))), flush=True)
    assert np.all(np.isfinite(u)) and u[:, 0].min() >= 0.0 and np.all(dts > 0)
    u2, t2, dts2 = _run(p, 200)
    assert same_bits(u, u2) and same_bits(dts, dts2) and t == t2


def test_closed_basin_16m_keeps_its_volume(bumps_unstructured):
    """The 16.7M-node Poisson-disk mesh closed by walls on all four sides, the bumps scaled to stay
    below 0.7 and a Gaussian hump of 0.05 on the free surface of level 1 (wet everywhere): 200 fct steps
    keep sum_i m_i h_i to 1e-14 relative (the reference's fct step conserves the volume of a closed wet
    domain; its first-order step does not, see test_mass_balance.py), with depth >= 0 and finite."""
    import dataclasses
    import math
    from swe_amd import problems
    q = bumps_unstructured
    z = q.bath * (0.7 / q.bath.max())
    rr = q.mesh.rr
    eta = 1.0 + 0.05 * np.exp(-((rr[:, 0] - 400.0) ** 2 + (rr[:, 1] - 600.0) ** 2) / (2.0 * 50.0 ** 2))
    u0 = np.zeros((q.np, 3))
    u0[:, 0] = eta - z
    params = dict(q.params, alpha_limit=False)
    p = dataclasses.replace(q, name="closed_basin", bath=z, u0=u0, params=params, h_sides=(), ux_sides=(2, 4),
                            uy_sides=(1, 3), bc=(problems.BC_WALLS, 0.0, 1.0, 0.0))
    v0 = math.fsum(p.lumped * u0[:, 0])
    u, t, dts = _run(p, 200)
    v = math.fsum(p.lumped * u[:, 0])
    print("closed basin 16.7M, 200 fct steps: t = %.6g, relative volume change %.3g" % (t, (v - v0) / v0), flush=True)
    assert np.all(np.isfinite(u)) and u[:, 0].min() >= 0.0 and np.all(dts > 0)
    assert abs(v - v0) <= 1e-14 * v0
