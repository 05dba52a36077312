"""The P1 set-up on the GPU (swe_assemble_device): the reference's construct_matrices output on
its own meshes (fixtures written by the reference build), and the host assembly's bits on larger
generated meshes, including the consistent mass and the diagonal positions."""
import numpy as np
import pytest

from conftest import golden, same_bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["par41_modes.npz", "dambreak_run.npz", "cav025_run.npz", "cdisk_coriolis.npz",
                                  "bumps_run.npz", "par_friction.npz"])
def test_device_assembly_is_the_references(name):
    from swe_amd.solver import assemble
    z = golden(name)
    print(name, flush=True)
    ia, ja, diag, mass, lumped, cij = assemble(z["jj"], z["rr"], device=True)
    assert same_bits(ia, z["ia"]) and same_bits(ja, z["ja"])
    assert same_bits(lumped, z["lumped"]) and same_bits(cij, z["cij"])
    assert np.all(ja[diag] == np.arange(ia.size - 1))


@pytest.mark.parametrize("n_side,seed", [(129, 3), (700, 11)])
def test_device_assembly_equals_host_assembly(n_side, seed):
    from swe_amd import mesh as M
    from swe_amd.solver import assemble
    m = M.renumber_first_touch(M.jittered_delaunay_mesh(n_side, 1000.0, jitter=0.25, seed=seed))
    print("mesh", m.np, flush=True)
    host = assemble(m.jj, m.rr)
    dev = assemble(m.jj, m.rr, device=True)
    for a, b in zip(host, dev):
        a, b = np.asarray(a), np.asarray(b)
        assert a.shape == b.shape and a.dtype == b.dtype
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))     # every bit, signs of zeros too


def test_device_assembly_rejects_large_fans_and_bad_input():
    from swe_amd.solver import SweError, assemble
    # one node shared by 40 elements: beyond the device rows' bound, reported, not assembled
    k = 40
    ang = 2 * np.pi * np.arange(k) / k
    rr = np.vstack([[0.0, 0.0], np.stack([np.cos(ang), np.sin(ang)], 1)])
    jj = np.array([[0, 1 + q, 1 + (q + 1) % k] for q in range(k)])
    with pytest.raises(SweError):
        assemble(jj, rr, device=True)
    with pytest.raises(SweError):
        assemble(np.array([[0, 1, 7]]), np.zeros((3, 2)), device=True)



def _scrambled(p, seed):
    """The case `p` with its nodes renumbered at random (neighbours far apart in the numbering)."""
    q = np.random.default_rng(seed).permutation(p.np)      # new number of each node
    inv = np.argsort(q)
    lists = [q[np.asarray(x, np.int64)] for x in p.dirichlet()]
    return q[p.mesh.jj], p.mesh.rr[inv], p.bath[inv], p.u0[inv], lists


@pytest.mark.parametrize("case", ["paraboloid", "dam_break", "bumps", "bumps_scrambled"])
def test_solver_from_device_assembly_steps_like_host_solver(case):
    """swe_create_from_assembly: the layout built on the GPU from the device assembly gives every
    bit of the states and time steps of swe_create on the host assembly (inflow, walls, friction,
    FCT with the alpha limiter, first-order viscosity; a scrambled numbering with far columns)."""
    from swe_amd import problems as PR
    from swe_amd.solver import Solver, assemble
    n = 20
    if case == "paraboloid":
        p = PR.paraboloid(n=41)
    elif case == "dam_break":
        p = PR.dam_break(nx=201, ny=41)
    elif case == "bumps":
        p = PR.bumps(n_side=129)
    else:
        p = PR.bumps(n_side=257)        # 66,049 nodes: scrambled columns beyond 16-bit offsets
    print(case, p.np, flush=True)
    if case == "bumps_scrambled":
        jj, rr, bath, u0, lists = _scrambled(p, 5)
        ia, ja, _diag, _mass, lumped, cij = assemble(jj, rr)
        kw = dict(h_dir=lists[0], ux_dir=lists[1], uy_dir=lists[2], **p.params)
        host = Solver(ia, ja, cij, lumped, bath, **kw)
        dev = Solver.from_elements(jj, rr, bath, **kw)
        for s in (host, dev):
            s.set_bc(*p.bc)
    else:
        u0 = p.u0
        host = p.solver()
        dev = p.solver(device_setup=True)
    out = []
    for s in (host, dev):
        s.set_state(u0)
        t, dts = s.run(n, time=0.0)
        out.append((s.get_state(), t, dts))
        s.close()
    (a, ta, da), (b, tb, db) = out
    assert same_bits(a, b) and same_bits(da, db) and ta == tb


def test_solver_from_device_assembly_rejects_bad_input():
    from swe_amd import problems as PR
    from swe_amd.solver import Solver, SweError
    p = PR.paraboloid(n=11)
    with pytest.raises(SweError):
        Solver.from_elements(p.mesh.jj, p.mesh.rr, p.bath, h_dir=[p.np], **p.params)    # node out of range
    with pytest.raises(SweError):
        Solver.from_elements(p.mesh.jj, p.mesh.rr, p.bath[:-1], **p.params)             # bath too short


@pytest.mark.parametrize("n", [2, 3, 9])
def test_solver_from_device_assembly_on_tiny_meshes(n):
    """Meshes of 4, 9 and 81 nodes (one partial 64-row slice, rows of 3 to 7 entries): the device
    layout and the host layout give the same bits over 10 first-order paraboloid steps."""
    from swe_amd import mesh as M
    from swe_amd import problems as PR
    from swe_amd.solver import Solver, assemble
    m = M.renumber_first_touch(M.structured_mesh(n, n, 0.0, 4.0, 0.0, 4.0))
    htiny = 4.0e-4
    bath = -0.1 * (1.0 - ((m.rr[:, 0] - 2.0) ** 2 + (m.rr[:, 1] - 2.0) ** 2))
    u0 = PR.paraboloid_state(m.rr, 0.0, htiny)
    kw = dict(gravity=9.81, htiny=htiny, cfl=0.25, max_water_h=4.0, viscosity="viscous", type_test=5)
    ia, ja, _diag, _mass, lumped, cij = assemble(m.jj, m.rr)
    out = []
    for s in (Solver(ia, ja, cij, lumped, bath, **kw), Solver.from_elements(m.jj, m.rr, bath, **kw)):
        s.set_state(u0)
        t, dts = s.run(10, time=0.0)
        out.append((s.get_state(), t, dts))
        s.close()
    assert same_bits(out[0][0], out[1][0]) and same_bits(out[0][2], out[1][2]) and out[0][1] == out[1][1]


def test_device_locality_order_forced(monkeypatch):
    """SWE_REORDER=2 forces the Hilbert-curve order of the device layout (and RCM on the host) on a mesh
    whose numbering is already good: the results keep every bit."""
    from swe_amd import problems as PR
    p = PR.bumps(n_side=129)
    out = []
    for mode, dev in (("0", False), ("2", True), ("2", False)):
        monkeypatch.setenv("SWE_REORDER", mode)
        with p.solver(device_setup=dev) as s:
            s.set_state(p.u0)
            t, dts = s.run(20, time=0.0)
            out.append((s.get_state(), t, dts))
    for u, t, d in out[1:]:
        assert same_bits(u, out[0][0]) and same_bits(d, out[0][2]) and t == out[0][1]
