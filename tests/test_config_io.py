"""ICARAT's TOML configuration files and legacy-VTK result files
(icarat_hip_create_from_toml, icarat_hip_write_vtk): host-side checks."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_toml_cantilever_equals_api_built_problem(capi):
    c, info = capi.Context.from_toml(os.path.join(GOLDEN, "cantilever_hexa8.toml"))
    assert (info.nx, info.ny, info.nz) == (12, 4, 4)
    assert (info.lx, info.ly, info.lz) == (12.0, 4.0, 4.0)
    assert info.designs_0 == 0.3 and info.filter_radius == 1.5
    assert (info.n_dirichlet, info.n_neumann) == (1, 5)
    ref = capi.cantilever(12, 4, 4)
    assert (c.numnp, c.nelx, c.numeq) == (ref.numnp, ref.nelx, ref.numeq)
    assert np.array_equal(c.dof(), ref.dof())
    assert np.array_equal(c.fext(), ref.fext())


def test_toml_face_load_and_box_voxels(capi, tmp_path):
    p = tmp_path / "beam.toml"
    p.write_text('''
[mesh]
dimension = 3
type = "hexa8"
division = [10, 4, 2]     # comment after a value
length = [5.0, 2.0, 2.0]

[boundary_condition]
dirichlet = [
  [[1, 1, 1, 0, 0, 0], [0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
   [[0.0, 0.0], [0.0, 2.0], [0.0, 2.0]]],
]
neumann = [[[0.0, 0.0, -1.0e-3], [[5.0, 5.0], [0.0, 2.0], [0.0, 2.0]]]]

[material]
young = [0.001, 100.0]
poisson = 0.3
penalty = 3.0
''')
    c, info = capi.Context.from_toml(str(p))
    ref = capi.Context(10, 4, 2, lx=5.0, ly=2.0, lz=2.0, young1=0.001, young2=100.0)
    ref.add_dirichlet([1, 1, 1, 0, 0, 0], [0.0] * 6, [[0, 0], [0, 2], [0, 2]])
    ref.make_dof()
    ref.add_neumann([0.0, 0.0, -1e-3], [[5, 5], [0, 2], [0, 2]])
    assert info.designs_0 == 1.0 and info.filter_radius == 0.0
    assert np.array_equal(c.fext(), ref.fext())
    # a face load of val over the face: the nodal loads sum to val x face area
    assert abs(c.fext().sum() - (-1e-3 * 2.0 * 2.0)) < 1e-15
    assert np.array_equal(c.ke0(), ref.ke0())


@pytest.mark.parametrize("text,msg", [
    ('[mesh]\ndimension = 2\ntype = "quad4"\n', "3-D"),
    ('[mesh]\ndimension = 3\ntype = "tetra4"\n', "hexa8"),
    ('[mesh]\ndimension = 3\ntype = "hexa8"\ndivision = [2, 2, 2]\n', "length"),
    ('[mesh]\ndimension = 3\ntype = "hexa8"\ndivision = [2, 2, 2\n', "line"),
    ('[mesh]\nx = {a = 1}\n', "inline"),
])
def test_toml_errors(capi, tmp_path, text, msg):
    p = tmp_path / "bad.toml"
    p.write_text(text)
    with pytest.raises(capi.IcaratError) as e:
        capi.Context.from_toml(str(p))
    assert msg in str(e.value)


def _read_vtk(path):
    lines = open(path).read().split("\n")
    sec = {}
    i = 0
    while i < len(lines):
        w = lines[i].split()
        if w and w[0] in ("POINTS", "CELLS", "CELL_TYPES", "POINT_DATA", "CELL_DATA"):
            sec[w[0]] = (i, w)
        i += 1
    return lines, sec


def test_vtk_mesh_and_density(capi, oracle, tmp_path):
    """The reference writer's layout (io/export_vtk.hpp:47-190): header, float
    points, hexa8 cells in the reference's local node order, cell type 12,
    CELL_DATA density with default stream formatting."""
    c = capi.Context(3, 2, 2, lx=1.5, ly=2.0, lz=1.0)
    c.make_dof()
    x = np.linspace(0.01, 1.0, c.nelx)
    path = str(tmp_path / "res1.vtk")
    c.write_vtk(path, x)
    lines, sec = _read_vtk(path)
    assert lines[:4] == ["# vtk DataFile Version 4.1", "vtk output", "ASCII",
                         "DATASET UNSTRUCTURED_GRID"]
    i, w = sec["POINTS"]
    assert w == ["POINTS", str(c.numnp), "float"]
    pts = np.array([[float(v) for v in lines[i + 1 + n].split()] for n in range(c.numnp)])
    o = oracle.Problem(3, 2, 2, 1.5, 2.0, 1.0)
    assert np.allclose(pts, o.coords(), rtol=0, atol=1e-12)
    i, w = sec["CELLS"]
    assert w == ["CELLS", str(c.nelx), str(9 * c.nelx)]
    cells = np.array([[int(v) for v in lines[i + 1 + e].split()] for e in range(c.nelx)])
    assert np.all(cells[:, 0] == 8)
    assert np.array_equal(cells[:, 1:], o.connectivity())
    i, w = sec["CELL_TYPES"]
    assert all(lines[i + 1 + e] == "12" for e in range(c.nelx))
    i, w = sec["CELL_DATA"]
    assert lines[i + 1] == "SCALARS density float"
    assert lines[i + 2] == "LOOKUP_TABLE default"
    dens = np.array([float(lines[i + 3 + e]) for e in range(c.nelx)])
    assert np.allclose(dens, x, rtol=1e-6, atol=0)
    assert lines[i + 3] == "%g" % x[0]
    assert "POINT_DATA" not in sec  # no solution yet
