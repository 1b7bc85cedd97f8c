"""A reference configuration file driven through the GPU loop, and the result
file with the solution fields (icarat_hip_create_from_toml, icarat_hip_write_vtk)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _section(lines, head):
    for i, l in enumerate(lines):
        if l.startswith(head):
            return i
    raise AssertionError(head)


def test_toml_driver_loop_and_vtk(gpu, tmp_path):
    c, info = gpu.Context.from_toml(os.path.join(GOLDEN, "cantilever_hexa8.toml"))
    ref = gpu.cantilever(12, 4, 4)
    for p in (c, ref):
        p.filter_setup(info.filter_radius)
    x0 = np.full(c.nelx, info.designs_0)
    xa, ha, _ = c.simp_run(x0, 3, tol=1e-10)
    xb, hb, _ = ref.simp_run(x0, 3, tol=1e-10)
    assert np.array_equal(xa, xb) and np.array_equal(ha, hb)
    # one solve at the final design, then the result file
    xf = c.density_filter(xa)
    c.set_density(xf)
    u, st = c.solve(tol=1e-10)
    m, _ = c.mises()
    path = str(tmp_path / "res3.vtk")
    c.write_vtk(path, xf)
    lines = open(path).read().split("\n")
    i = _section(lines, "POINT_DATA")
    assert lines[i + 1] == "VECTORS displacement float"
    disp = np.array([[float(v) for v in lines[i + 2 + n].split()] for n in range(c.numnp)])
    dof = c.dof().reshape(-1, 3)
    nodal = np.where(dof < c.numeq, u[np.minimum(dof, c.numeq - 1)], 0.0)
    scale = np.max(np.abs(nodal))
    assert np.max(np.abs(disp - nodal)) < 1e-5 * scale
    i = _section(lines, "CELL_DATA")
    assert lines[i + 1] == "SCALARS density float"
    j = _section(lines, "SCALARS mises")
    mv = np.array([float(lines[j + 2 + e]) for e in range(c.nelx)])
    assert np.max(np.abs(mv - m)) < 1e-5 * np.max(m)
