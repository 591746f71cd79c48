"""Legacy-VTK output (gf_write_vtk / write_vtk / printToVTK, replacing
include/mesh/vtk_writer.h:8-73) -- host only.  Known answer: the file the
reference writer produces for a 2 x 1 plate (its header lines, POINTS with
the type name "double64", CELLS with corners around the element, cell type
9, the u v w of each node as the vector field "disp", reals as the default
stream format prints them)."""
import numpy as np
import pytest

import gpu_fem_amd as G
from gpu_fem_amd import meshes as M

EXPECTED = """# vtk DataFile Version 2.0
TACS GPU shell writer
ASCII
DATASET UNSTRUCTURED_GRID
POINTS 6 double64
0 0 0
0.5 0 0
1 0 0
0 0.25 0
0.5 0.25 0
1 0.25 0
CELLS 2 10
4 0 1 4 3
4 1 2 5 4
CELL_TYPES 2
9
9
POINT_DATA 6
VECTORS disp double64
0 1 -2
1e-07 1.001 -1.23457e+06
2 1 -2
3 1 -2
4 1 -2
5 1 -2
"""


def test_plate_known_answer(tmp_path):
    m = M.plate(2, 1, 1.0, 0.25)
    x = np.zeros(6 * m.nnodes)
    x[0::6] = np.arange(m.nnodes)
    x[1::6] = 1.0
    x[2::6] = -2.0
    x[3::6] = 99.0  # rotations are not written
    x[6:9] = [1e-7, 1.001, -1234567.0]
    p = tmp_path / "plate.vtk"
    G.write_vtk(str(p), m, x)
    assert p.read_text() == EXPECTED


def test_without_displacements_and_bad_input(tmp_path):
    m = M.plate(2, 1, 1.0, 0.25)
    p = tmp_path / "zero.vtk"
    G.write_vtk(str(p), m)
    assert p.read_text().endswith("VECTORS disp double64\n" + "0 0 0\n" * 6)
    with pytest.raises(ValueError):
        G.write_vtk(str(p), m, np.zeros(5))
    bad = M.plate(2, 1)
    bad.conn = bad.conn.copy()
    bad.conn[0, 0] = 99
    with pytest.raises(G.GfError):
        G.write_vtk(str(p), bad)
