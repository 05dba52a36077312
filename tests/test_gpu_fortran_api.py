"""The Fortran interface blocks of the device set-up and of compute_errors, called from a Fortran
program (tests/fortran/api_check.f90, built by `make -C tests/fortran`): its final state, time, every dt
and the 12 error sums (and the 4 values of the unit-10 diagnostic) equal, bit for bit, those of the
same calls made from Python (which other tests pin against the oracle)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu

API_CHECK = os.path.join(REPO, "tests", "fortran", "_build", "api_check")


def _write_input(path, p, nsteps):
    from swe_amd.solver import VISCOSITY
    pr = p.params
    hd, uxd, uyd = [np.asarray(x, np.int32) + 1 for x in p.dirichlet()]
    kind, amp, period, level = p.bc
    vt = VISCOSITY[pr["viscosity"]] if isinstance(pr["viscosity"], str) else int(pr["viscosity"])
    with open(path, "wb") as f:
        f.write(np.array([p.np, p.mesh.jj.shape[0], hd.size, uxd.size, uyd.size, nsteps, kind], "<i4").tobytes())
        f.write(np.array([amp, period, level, pr["gravity"], pr["htiny"], pr["cfl"], pr["max_water_h"],
                          pr.get("mannings", 0.0)], "<f8").tobytes())
        f.write(np.array([pr["type_test"], vt, pr.get("viscous_type", 2), int(bool(pr.get("alpha_limit", False))),
                          1], "<i4").tobytes())
        f.write(np.ascontiguousarray(p.mesh.jj + 1, "<i4").tobytes())
        f.write(np.ascontiguousarray(p.mesh.rr, "<f8").tobytes())
        f.write(np.ascontiguousarray(p.bath, "<f8").tobytes())
        f.write(np.ascontiguousarray(p.u0, "<f8").tobytes())
        for x in (hd, uxd, uyd):
            f.write(np.ascontiguousarray(x, "<i4").tobytes())


@pytest.mark.skipif(not os.path.exists(API_CHECK), reason="tests/fortran was not built (make -C tests/fortran)")
@pytest.mark.parametrize("case", ["paraboloid", "bumps"])
def test_fortran_calls_equal_python_calls(tmp_path, case):
    from swe_amd import problems as PR
    p = PR.paraboloid(n=41) if case == "paraboloid" else PR.bumps(n_side=65, L=1000.0, seed=11)
    n = 25
    _write_input(str(tmp_path / "in.bin"), p, n)
    print("api_check", case, p.np, flush=True)
    r = subprocess.run([API_CHECK], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    with p.solver(device_setup=True) as s:
        s.set_state(p.u0)
        t, dts = s.run(n, time=0.0)
        u = s.get_state()
        # test 5 is the paraboloid's own analytic solution; the bumps case (type_test 4) is compared
        # with test 4's dam-break solution, which is just as good a check of the calls
        ev = s.error_evaluator(p.mesh.jj, p.mesh.rr)
        errs = ev.sums(t).reshape(-1)
        flow = ev.ns_l1_ratio()
        ev.close()
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = np.fromfile(str(tmp_path / "out.bin"), "<f8")
    npn = p.np
    m = 3 * npn + 1 + n + 12
    fu, ft, fd, fe = out[:3 * npn].reshape(npn, 3), out[3 * npn], out[3 * npn + 1:3 * npn + 1 + n], out[m - 12:m]
    assert np.array_equal(fu, u) and ft == t and np.array_equal(fd, dts)
    assert np.array_equal(fe, errs)
    assert "Relative L1 error on h" in r.stdout
    # the batch part: two instances of the same mesh, each the solver's bits
    rest = out[m:]
    ub = rest[:6 * npn].reshape(2, npn, 3)
    tb, nst, eb = rest[6 * npn:6 * npn + 2], rest[6 * npn + 2:6 * npn + 4], rest[6 * npn + 4:6 * npn + 16]
    assert eb.size == 12 and rest.size == 6 * npn + 20
    assert np.array_equal(rest[6 * npn + 16:], flow)
    assert np.array_equal(ub[0], u) and np.array_equal(ub[1], u)
    assert np.all(tb == t) and np.all(nst == n)
    assert np.array_equal(eb, errs)
