"""The C-ABI library without a GPU: it loads, exports every declared symbol, its host-side set-up
is the reference's, and malformed inputs are rejected before any device work."""
import os
import re

import numpy as np
import pytest

from conftest import REPO, golden, same_bits

HEADER = os.path.join(REPO, "include", "swe_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(swe_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from swe_amd import solver
    lib = solver.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(solver.SIGNATURES)


@pytest.mark.parametrize("name", ["par41_modes.npz", "dambreak_run.npz", "bumps_run.npz", "cav025_run.npz",
                                  "cdisk_coriolis.npz"])
def test_assembly_is_the_references(name):
    from swe_amd.solver import assemble
    z = golden(name)
    ia, ja, diag, mass, lumped, cij = assemble(z["jj"], z["rr"])
    assert same_bits(ia, z["ia"]) and same_bits(ja, z["ja"])
    assert same_bits(lumped, z["lumped"]) and same_bits(cij, z["cij"])
    assert np.all(ja[diag] == np.arange(ia.size - 1))


def test_problem_generators_reproduce_reference_inputs():
    from swe_amd import problems
    z = golden("par41_modes.npz")
    p = problems.paraboloid(41)
    for k in ("ia", "ja", "lumped", "cij", "bath", "u0"):
        assert same_bits(getattr(p, k), z[k]), k
    z = golden("dambreak_run.npz")
    p = problems.dam_break(101, 21, seed=7)
    for k in ("ia", "ja", "lumped", "cij", "bath", "u0"):
        assert same_bits(getattr(p, k), z[k]), k
    hd, uxd, uyd = p.dirichlet()
    assert same_bits(uxd, z["ux_dir"]) and same_bits(uyd, z["uy_dir"]) and hd.size == 0
    z = golden("bumps_run.npz")
    p = problems.bumps(n_side=65, L=1000.0, seed=11)
    for k in ("ia", "ja", "lumped", "cij", "bath", "u0"):
        assert same_bits(getattr(p, k), z[k]), k
    hd, uxd, uyd = p.dirichlet()
    assert same_bits(hd, z["h_dir"]) and same_bits(uxd, z["ux_dir"]) and same_bits(uyd, z["uy_dir"])


def _desc_args(z):
    return dict(gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0)


def test_create_rejects_malformed_input_before_device_work():
    from swe_amd.solver import SWE_EBADOPT, SWE_EINVAL, Solver, SweError
    z = golden("cav025_run.npz")
    ok = dict(gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0)
    bad_ja = z["ja"].copy()
    bad_ja[3] = z["ia"].size + 5          # column out of range
    with pytest.raises(SweError) as e:
        Solver(z["ia"], bad_ja, z["cij"], z["lumped"], z["bath"], **ok)
    assert e.value.status == SWE_EINVAL
    bad_m = z["lumped"].copy()
    bad_m[0] = 0.0
    with pytest.raises(SweError) as e:
        Solver(z["ia"], z["ja"], z["cij"], bad_m, z["bath"], **ok)
    assert e.value.status == SWE_EINVAL
    with pytest.raises(SweError) as e:
        Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], h_dir=[z["ia"].size + 3], **ok)
    assert e.value.status == SWE_EINVAL
    # FCT is only programmed with type2 (SOURCES/update.f90:272-275)
    with pytest.raises(SweError) as e:
        Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], viscosity="fct", viscous_type=1, **ok)
    assert e.value.status == SWE_EBADOPT


@pytest.mark.parametrize("tt,viscosity,ux", [(0, "viscous", ()), (14, "fct", ()), (-3, "viscous", ()),
                                             (13, "fct", ()), (13, "entropy_visc", ()), (13, "viscous", (0, 1))])
def test_create_rejects_what_the_reference_leaves_undefined(tt, viscosity, ux):
    """Refused with SWE_EBADOPT before any device work (batch admission checks its parameters with
    the same functions): type_test outside 1..13, where compute_lambda_vacc STOPs with 'BUG in
    compute_lambda' (SOURCES/condlim.f90:897,952-954); the SGN model (test 13) with entropy viscosity
    or FCT, whose entropy_residual reads Entprime(4:5,:) that the reference never sets
    (update.f90:853-862); and SGN Dirichlet nodes for h*eta, h*w, whose sol_anal the reference
    defines only at t <= 1e-10 (condlim.f90:801-826)."""
    from swe_amd.solver import SWE_EBADOPT, Solver, SweError
    z = golden("cav025_run.npz")
    ok = dict(gravity=9.81, htiny=1e-4, cfl=0.25, max_water_h=1.0)
    with pytest.raises(SweError) as e:
        Solver(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], type_test=tt, viscosity=viscosity, ux_dir=ux, **ok)
    assert e.value.status == SWE_EBADOPT
    assert ("compute_lambda" in str(e.value)) == (tt != 13)


def test_assemble_rejects_bad_elements():
    from swe_amd.solver import SweError, assemble
    with pytest.raises(SweError):
        assemble(np.array([[0, 1, 7]]), np.zeros((3, 2)))
