"""Copies data files the reference holds, for the tests: the first 401 lines of PARABOLOID/h.plt
(plot_scalar_field output: header and 99 elements) and SOURCES/BACKUP/h_1.vtk (vtk_2d output of a
508-node mesh) for the format tests of swe_amd/output.py, and the mesh
MESHES/SGN_CANAL/rectangle.3.FEM (365 nodes) that the SGN drop-in test runs the reference program
on.  Run where the reference sources are (they are not needed afterwards):

    python tests/golden/extract_output_fixtures.py /root/reference
"""
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(ref, "PARABOLOID", "h.plt")) as f:
    lines = f.read().split("\n")[:401]
with open(os.path.join(here, "ref_h_plt_head.txt"), "w") as f:
    f.write("\n".join(lines) + "\n")
with open(os.path.join(ref, "SOURCES", "BACKUP", "h_1.vtk")) as f:
    vtk = f.read()
with open(os.path.join(here, "ref_h_1_vtk.txt"), "w") as f:
    f.write(vtk)
with open(os.path.join(ref, "MESHES", "SGN_CANAL", "rectangle.3.FEM"), "rb") as f:
    fem = f.read()
with open(os.path.join(here, "sgn_rectangle3.FEM"), "wb") as f:
    f.write(fem)
