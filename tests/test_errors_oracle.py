"""compute_errors (SOURCES/main.f90:198-263) in the oracle: pinned bitwise against the reference's own
sol_anal, ns_anal_l1 and ns_anal_0 (tests/golden/errors.npz, written by the reference build), and its
sequential sums against the exact sums of the same summands (math.fsum) within the rigorous bound of
recursive summation, |S_seq - S| <= (n-1)u/(1-(n-1)u) * sum|t_i|."""
import math

import numpy as np
import pytest

from conftest import golden


def _case(key):
    name, st, t = key.split("|")
    z = golden(name + ".npz")
    u = z["states"][0] if st == "states0" else z["states"][1] if st == "states1" else z["final"]
    t = float(z["time"]) if t == "time" else float(t)
    return z, u, t


def _keys():
    return sorted(golden("errors.npz"))


@pytest.mark.parametrize("key", _keys())
def test_oracle_errors_are_the_references(oracle, key):
    z, u, t = _case(key)
    g, htiny = z["scalars"][0], z["scalars"][1]
    out = oracle.compute_errors(z["jj"], z["rr"], u, int(z["type_test"]), t, g, htiny)
    assert np.array_equal(out, golden("errors.npz")[key])


@pytest.mark.parametrize("key", _keys())
def test_oracle_error_sums_within_the_summation_bound(oracle, key):
    z, u, t = _case(key)
    g, htiny = z["scalars"][0], z["scalars"][1]
    out, terms = oracle.compute_errors(z["jj"], z["rr"], u, int(z["type_test"]), t, g, htiny, terms=True)
    n = terms.shape[-1]
    gam = (n - 1) * 2.0 ** -53 / (1 - (n - 1) * 2.0 ** -53)
    for k in range(3):
        for q in range(4):
            exact = math.fsum(terms[k, q])
            seq = out[k, q] ** 2 if q >= 2 else out[k, q]
            # the L2 entries are square roots: compare the sums (sqrt is correctly rounded, so
            # squaring it back is within 2 ulp of the sequential sum)
            slack = 4 * 2.0 ** -53 * abs(seq) if q >= 2 else 0.0
            assert abs(seq - exact) <= gam * np.abs(terms[k, q]).sum() + slack


def test_sol_anal_of_test_5_is_the_initial_state(oracle):
    """sol_anal(k, rr, 0) of test 5 at the mesh nodes is the reference's initial state (init, the
    run fixture's u0), evaluated there by the reference itself."""
    z = golden("par41_modes.npz")
    g, htiny = z["scalars"][0], z["scalars"][1]
    rr = z["rr"]
    u = np.array([[oracle.sol_anal(5, k, x, y, 0.0, g, htiny) for k in (1, 2, 3)] for x, y in rr])
    assert np.array_equal(u, z["u0"])


def test_error_of_the_exact_field_vanishes(oracle):
    """A plain invariant: with u set to the exact solution where it is constant (test 4 for
    x < x0 - t*speed: h = hL, q = 0 on all of [0,4]^2) the errors are the rounding of the P1
    interpolation (a few ulp of the norm), the norms are hL * area in L1 and hL * sqrt(area) in L2."""
    from swe_amd import mesh as M
    m = M.structured_mesh(9, 9, 0.0, 4.0, 0.0, 4.0)
    u = np.zeros((m.np, 3))
    u[:, 0] = 0.005
    out = oracle.compute_errors(m.jj, m.rr, u, 4, 1.0, 9.81, 1e-7)
    assert out[0, 0] <= 8 * 2.0 ** -53 * out[0, 1] and out[0, 2] <= 8 * 2.0 ** -53 * out[0, 3]
    assert abs(out[0, 1] - 0.005 * 16.0) <= 1e-15 and abs(out[0, 3] - 0.005 * 4.0) <= 1e-15
    assert np.all(out[1:, :] == 0.0)


@pytest.mark.parametrize("test", [3, 6, 7])
def test_oracle_bump_flow_errors_are_the_references(oracle, test):
    """Tests 3, 6, 7 (steady flows over the bump: roots of a cubic per point) at t = 0 and t = 1."""
    z = golden("errors_bump.npz")
    g, htiny = z["scalars_%d" % test]
    for t in (0, 1):
        out = oracle.compute_errors(z["jj"], z["rr"], z["u0_%d" % test], test, float(t), g, htiny)
        assert np.array_equal(out, z["errors_%d_t%d" % (test, t)]), t


def _flow_case(key):
    name, st = key.split("|")
    z = golden(name + ".npz")
    if st.startswith("snap"):
        u = z["snaps"][list(z["snap_steps"]).index(int(st[4:]))]
    else:
        u = z[st]
    return z, u


@pytest.mark.parametrize("key", sorted(golden("flow.npz")))
def test_oracle_flow_ratio_is_the_references(oracle, key):
    """The time loop's unit-10 diagnostic of tests 1, 2, 10 (main.f90:94-99): the ns_l1 norms of qx, qy
    and h*sqrt(h) and their ratio, bit for bit the reference's (tests/golden/flow.npz)."""
    z, u = _flow_case(key)
    out = oracle.ns_l1_ratio(z["jj"], z["rr"], u, float(z["scalars"][0]))
    assert np.array_equal(out, golden("flow.npz")[key])


@pytest.mark.parametrize("key", sorted(golden("flow.npz")))
def test_oracle_flow_sums_within_the_summation_bound(oracle, key):
    z, u = _flow_case(key)
    out, terms = oracle.ns_l1_ratio(z["jj"], z["rr"], u, float(z["scalars"][0]), terms=True)
    n = terms.shape[-1]
    gam = (n - 1) * 2.0 ** -53 / (1 - (n - 1) * 2.0 ** -53)
    for f in range(3):
        assert np.all(terms[f] >= 0.0)
        assert abs(out[f] - math.fsum(terms[f])) <= gam * terms[f].sum()


def test_flow_ratio_of_uniform_fields(oracle):
    """Closed form: for uniform h = H, qx = a, qy = b on a jittered [0,100]x[0,20] mesh the three norms
    are |a|, |b| and H^1.5 times the area 2000 (P1 interpolation of a constant is exact), and the
    ratio is 0.5 (|a| + |b|) / (g H^1.5).  A dropped absolute value (b < 0) or a swapped component
    fails it."""
    from swe_amd import mesh as M
    m = M.structured_mesh(41, 9, 0.0, 100.0, 0.0, 20.0, jitter=0.2, seed=3)
    H, a, b, g = 0.64, 0.3, -0.05, 9.81
    u = np.empty((m.np, 3))
    u[:, 0], u[:, 1], u[:, 2] = H, a, b
    out = oracle.ns_l1_ratio(m.jj, m.rr, u, g)
    area = 2000.0
    assert out[0] == pytest.approx(abs(a) * area, rel=1e-13)
    assert out[1] == pytest.approx(abs(b) * area, rel=1e-13)
    assert out[2] == pytest.approx(H ** 1.5 * area, rel=1e-13)
    assert out[3] == pytest.approx(0.5 * (abs(a) + abs(b)) / (g * H ** 1.5), rel=1e-13)


def test_oracle_unit10_series_is_the_reference_programs(oracle):
    """The reference program itself (main.f90) on test 1, FCT + alpha, 268 steps to Tfinal = 1: the
    oracle stepped one step at a time from the reference's set-up reaches every step's time and writes
    every unit-10 ratio (orc_ns_l1_ratio) bit for bit as the program did (tests/golden/unit10_test1.npz)."""
    z = golden("unit10_test1.npz")
    g, htiny, hmax, cfl = z["scalars"]
    o = oracle.Oracle(z["ia"], z["ja"], z["cij"], z["lumped"], z["bath"], gravity=g, htiny=htiny, cfl=cfl,
                      max_water_h=hmax, viscosity_type=oracle.FCT, alpha_limit=True, type_test=1)
    try:
        u, t = z["u0"], float(z["time0"])
        for tr, ratio in zip(z["times"], z["ratios"]):
            u, _, t, err = o.run(u, 1, t0=t)
            assert err == 0 and t == tr
            assert oracle.ns_l1_ratio(z["jj"], z["rr"], u, g)[3] == ratio
    finally:
        o.close()
