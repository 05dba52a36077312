"""Field output in the reference's formats: the reference's own output files (tests/golden/ref_*,
copied by tests/golden/extract_output_fixtures.py) are re-emitted byte for byte from the values they
hold, which pins the layout of every number, header and record."""
import os

import numpy as np

from conftest import GOLDEN


def _read(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def test_plt_layout_is_the_references(tmp_path):
    from swe_amd.output import write_plt
    text = _read("ref_h_plt_head.txt")
    lines = text.split("\n")
    assert lines[:5] == ["$ DATA = CONTCURVE", "% contstyle = 2", "% nsteps = 50", "% meshplot  = false", ""]
    body = lines[5:]
    tris, vals = [], {}
    for k in range(0, len(body) - 1, 4):
        tri = []
        for l in body[k:k + 3]:
            x, y, u, n = l.split()
            vals[int(n) - 1] = (float(x), float(y), float(u))
            tri.append(int(n) - 1)
        assert body[k + 3] == ""
        tris.append(tri)
    npn = max(vals) + 1
    rr = np.zeros((npn, 2))
    u = np.zeros(npn)
    for n, (x, y, v) in vals.items():
        rr[n] = (x, y)
        u[n] = v
    out = tmp_path / "h.plt"
    write_plt(out, np.array(tris), rr, u)
    assert out.read_text() == text


def test_vtk_layout_is_the_references(tmp_path):
    from swe_amd.output import write_vtk
    text = _read("ref_h_1_vtk.txt")
    lines = text.split("\n")
    npn = int(lines[4].split()[1])
    rr = np.array([[float(v) for v in l.split()[:2]] for l in lines[5:5 + npn]])
    me = int(lines[5 + npn].split()[1])
    jj = np.array([[int(v) for v in l.split()[1:]] for l in lines[6 + npn:6 + npn + me]])
    k = 6 + npn + me + 1 + me
    assert lines[k].startswith("POINT_DATA")
    field = np.array([float(l) for l in lines[k + 3:k + 3 + npn]])
    out = tmp_path / "h_1.vtk"
    write_vtk(out, jj, rr, field, title="h_1.vtk", title_len=len(lines[1]) - 4)
    assert out.read_text() == text


def test_fortran_e_edit_descriptor():
    from swe_amd.output import fortran_e, list_directed_real
    assert fortran_e(0.0) == "0.0000000E+00"
    assert fortran_e(1.5) == "0.1500000E+01"
    assert fortran_e(-1.5) == "-.1500000E+01"          # no room for the optional leading zero
    assert fortran_e(0.999999996) == "0.1000000E+01"   # rounds into the next decade
    assert fortran_e(1.0e-120) == "0.1000000-119"      # three-digit exponents drop the E
    assert fortran_e(123456789.0) == "0.1234568E+09"
    assert list_directed_real(0.44711204761671658) == "  0.44711204761671658     "
    assert list_directed_real(1.1684898678821490e-152) == "   1.1684898678821490E-152"
