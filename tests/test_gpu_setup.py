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
