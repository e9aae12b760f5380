"""VTK output (DataWriter::writeVtk) of FE functions read from device memory:
the one-cell file against the text the reference's writer produces for it,
and on refined meshes (quadrilaterals, triangles, hexahedra; Q1, P1, Q2) the
points, the cell connectivity (each cell's points are its vertices in local
order) and the point values (the DOF values at the vertices)."""
import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm

pytestmark = pytest.mark.gpu


def parse_vtk(text):
    lines = text.split("\n")
    assert lines[0] == "# vtk DataFile Version 4.0" and lines[2] == "ASCII"
    assert lines[3] == "DATASET UNSTRUCTURED_GRID"
    n_pts = int(lines[4].split()[1])
    pts = np.array([[float(v) for v in lines[5 + i].split()] for i in range(n_pts)])
    k = 5 + n_pts
    assert lines[k] == ""
    hdr = lines[k + 1].split()
    assert hdr[0] == "CELLS"
    nc, total = int(hdr[1]), int(hdr[2])
    cells = [[int(v) for v in lines[k + 2 + c].split()] for c in range(nc)]
    assert sum(len(c) for c in cells) == total and all(c[0] == len(c) - 1 for c in cells)
    k = k + 2 + nc
    assert lines[k] == "" and lines[k + 1] == "CELL_TYPES %d" % nc
    types = [int(v) for v in lines[k + 2].split()]
    assert len(types) == nc
    k += 3
    assert lines[k] == "" and lines[k + 1] == "POINT_DATA %d" % n_pts
    k += 2
    data = {}
    while k < len(lines) and lines[k].startswith("SCALARS "):
        name = lines[k].split()[1]
        assert lines[k + 1] == "LOOKUP_TABLE default"
        data[name] = np.array([float(v) for v in lines[k + 2:k + 2 + n_pts]])
        k += 2 + n_pts
        assert lines[k] == "" and lines[k + 1] == ""
        k += 2
    return pts, [c[1:] for c in cells], types, data


def test_vtk_single_cell_text(ctx, tmp_path):
    d = pm.TDomain("UnitSquare", 0)
    s = pm.TFESpace(d, 0, 1, ctx)
    u = torch.tensor([0.5, -1.25, 3.0, 1e-20][:s.n_dof], dtype=torch.float64, device="cuda")
    path = str(tmp_path / "one.vtk")
    pm.write_vtk(s, {"c": u}, path)
    xyz = d.vertices(0)[0]
    dof = s.dof_coords()
    val = [float(u[np.argmin(np.sum((dof[:, :2] - x) ** 2, axis=1))]) for x in xyz]
    expect = ("# vtk DataFile Version 4.0\nfile created by ParMooN Time < 0 >\nASCII\n"
              "DATASET UNSTRUCTURED_GRID\nPOINTS 4 double\n"
              + "".join("%.8g %.8g 0\n" % (x[0], x[1]) for x in xyz)
              + "\nCELLS 1 5\n4 0 1 2 3 \n\nCELL_TYPES 1\n9 \n\nPOINT_DATA 4\n"
              "SCALARS c double\nLOOKUP_TABLE default\n" + "".join("%.12e\n" % v for v in val) + "\n\n\n")
    assert open(path).read() == expect


@pytest.mark.parametrize("geo,nref,order,vtype", [("UnitSquare", 3, 1, 9), ("TwoTriangles", 3, 1, 5),
                                                  ("UnitSquare", 2, 2, 9), ("Default_UnitCube_Hexa", 2, 2, 12)])
def test_vtk_refined_meshes(ctx, tmp_path, geo, nref, order, vtype):
    d = pm.TDomain(geo, nref)
    s = pm.TFESpace(d, nref, order, ctx)
    ex = "exp_sin_3d" if geo == "Default_UnitCube_Hexa" else "tcd_exp"
    u = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    pm.interpolate(ctx, s, pm.cd_params(ex, time=0.3), u)
    w = torch.arange(s.n_dof, dtype=torch.float64, device="cuda")
    path = str(tmp_path / "f.vtk")
    pm.write_vtk(s, {"u": u, "index": w}, path, time=0.3)
    text = open(path).read()
    assert text.split("\n")[1] == "file created by ParMooN Time < 0.3 >"
    pts, cells, types, data = parse_vtk(text)
    dim = 3 if geo == "Default_UnitCube_Hexa" else 2
    side = 2 ** nref + 1
    assert len(pts) == side ** dim and set(types) == {vtype}
    V = d.vertices(nref)
    assert len(cells) == len(V)
    for c, verts in zip(cells, V):
        assert np.allclose(pts[c][:, :dim], verts, atol=1e-8)
    dof = s.dof_coords()
    un = u.cpu().numpy()
    for p in range(len(pts)):
        g = int(round(data["index"][p]))
        assert np.allclose(dof[g, :dim], pts[p, :dim], atol=1e-8)
        assert abs(data["u"][p] - un[g]) <= 1e-12 * max(1.0, abs(un[g]))
