"""Field output in the reference's file formats, for frames and final fields of the device-resident
runs (the reference writes them from its host arrays after the time loop).

* ``write_plt``: the PLOTMTV contour file of ``plot_scalar_field``
  (SFEMaNS/FEMSUB/sub_plot.f90:703-778) for P1 meshes: four header lines and a blank line, then per
  element its three nodes as ``x y u n`` and a blank line.  Numbers are laid out as the reference's
  list-directed output gives them (its committed ``.plt`` files): each REAL(8) with 17 significant
  digits, F form in a 21-character field plus 5 blanks for 0.1 <= |x| < 1e17 and for 0, otherwise
  E form with a 3-digit exponent in a 26-character field; the node number in a 12-character field.
* ``write_vtk``: the legacy VTK file of ``vtk_2d`` (sub_plot.f90:1036-1095) for P1 meshes, with its
  edit descriptors: POINTS ``(2(E13.7,2X),A)``, CELLS ``(A,3(I8,1X))``, data ``(E13.7,2X)`` (trailing
  blanks of X editing are not written, as Fortran output does not extend a record with them).

Indices are 0-based here (``jj`` (me, 3)); the files carry the reference's numbering (1-based node
numbers in ``.plt``, 0-based point indices in VTK cells).
"""
from __future__ import annotations

import numpy as np

__all__ = ["write_plt", "write_vtk", "fortran_e", "list_directed_real"]


def list_directed_real(x: float) -> str:
    """One REAL(8) as the reference's list-directed WRITE lays it out (26 characters)."""
    x = float(x)
    if x == 0.0:
        return "0.0000000000000000".rjust(21) + " " * 5
    mant, exp = ("%.16e" % x).split("e")
    ex = int(exp)
    if 0.1 <= abs(x) < 1e17:
        dec = 16 - ex if ex >= 0 else 17
        return ("%.*f" % (dec, x)).rjust(21) + " " * 5
    return ("%sE%s%03d" % (mant, "-" if ex < 0 else "+", abs(ex))).rjust(26)


def fortran_e(x: float, w: int = 13, d: int = 7) -> str:
    """Fortran Ew.d output: 0.d...d E+ee, rounded to d significant digits; the optional leading
    zero is dropped when the field has no room for it (negative values at E13.7)."""
    x = float(x)
    if x == 0.0:
        body, ex = "0" * d, 0
    else:
        mant, exp = ("%.*e" % (d - 1, abs(x))).split("e")
        body, ex = mant.replace(".", ""), int(exp) + 1
    exps = ("E%s%02d" % ("-" if ex < 0 else "+", abs(ex))) if abs(ex) <= 99 else ("%s%03d" % ("-" if ex < 0 else "+", abs(ex)))
    s = ("-" if x < 0 else "") + "0." + body + exps
    if len(s) > w:
        s = s.replace("0.", ".", 1)
    return s.rjust(w)


def write_plt(path, jj, rr, u):
    """plot_scalar_field(jj, rr, uu, file_name) for a P1 mesh."""
    jj = np.asarray(jj)
    rr = np.asarray(rr, np.float64)
    u = np.asarray(u, np.float64)
    if jj.ndim != 2 or jj.shape[1] != 3:
        raise ValueError("write_plt handles P1 meshes (jj of shape (me, 3))")
    out = ["$ DATA = CONTCURVE", "% contstyle = 2", "% nsteps = 50", "% meshplot  = false", ""]
    for tri in jj:
        for n in tri:
            out.append(list_directed_real(rr[n, 0]) + list_directed_real(rr[n, 1]) + list_directed_real(u[n])
                       + ("%d" % (int(n) + 1)).rjust(12))
        out.append("")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_vtk(path, jj, rr, field, title=None, title_len=200):
    """vtk_2d(mesh, champ, unit_file, file_name) for a P1 mesh; the title line is 'vtk ' and the
    file name in a CHARACTER(title_len) variable, as the reference passes it."""
    jj = np.asarray(jj)
    rr = np.asarray(rr, np.float64)
    field = np.asarray(field, np.float64)
    if jj.ndim != 2 or jj.shape[1] != 3:
        raise ValueError("write_vtk handles P1 meshes (jj of shape (me, 3))")
    npn, me = rr.shape[0], jj.shape[0]
    name = title if title is not None else str(path).rsplit("/", 1)[-1]
    out = ["# vtk DataFile Version 3.0", "vtk " + name.ljust(title_len), "ASCII", "DATASET UNSTRUCTURED_GRID",
           "POINTS %7d float" % npn]
    out += [fortran_e(x) + "  " + fortran_e(y) + "   0.0 " for x, y in rr]
    out.append("CELLS %7d%8d" % (me, me * 4))
    out += ["3 %8d %8d %8d" % (a, b, c) for a, b, c in jj]
    out.append("CELL_TYPES %7d" % me)
    out += ["5"] * me
    out.append("POINT_DATA %7d" % npn)
    out += ["SCALARS scalars float 1", "LOOKUP_TABLE default"]
    out += [fortran_e(v) for v in field]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
