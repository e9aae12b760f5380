"""Writes tests/golden/hexa_numbering.json: the reference's GlobalNumbers of
the unit-cube hexahedral mesh after one and two regular refinements
(2x2x2 and 4x4x4 cells) for Q1 and Q2, with every input read from the
reference's own source files (run here, where the reference sources are
present; the JSON is committed data).

Derivation, step by step:

 1. Level-0 vertices.  TDomain::initialize_cube_hexa_mesh
    (src/Geometry/Tests.C:1815-1850) gives the 8 vertices as (comp, t, s) of
    the boundary description initializeDefaultCubeBdry (Tests.C:1767-1810):
    6 planes, each with the three vectors P, A, N.  TBdPlane::ReadIn
    (src/Geometry/BdPlane.C:116-132) sets B = A x N, and GetXYZofTS
    (BdPlane.C:46-54) maps (t, s) to P + t A + s B.
 2. Regular refinement.  Child k of a hexahedron has the vertices
    DatChildVertex[k] (src/Refinement/RefHexaRegDesc.C:32-36) of the 27
    vertices of the refined cell: 0-7 are the old vertices
    (DatNewVertexEqOldVertex, :226), vertex DatInteriorVertexOfEdge[e]
    (:257-259) is the midpoint of old edge e with the end points of
    THexahedron's DatEdgeVertex (src/Geometry/Hexahedron.C:21-23),
    DatInteriorVertexOfFace[f] (:264-265) is the point of old face f with
    the weights DatOldFaceNewVertexPos[f][8] (:231-245) of the face's
    vertices DatFaceVertex (Hexahedron.C:28-30), and
    DatInteriorVertexOfCell (:229) is the point DatPositionOfIntVert
    (:247-248) of the 8 old vertices.  The finest cells are visited depth
    first (children of cell 0 first, each in child order).  As a check the
    script also derives every new vertex as the centroid of the old vertices
    of the children that contain it (DatVertexChild, :38-42).
 3. Local DOFs.  Local DOF k of the continuous Q1 / Q2 hexahedron sits at the
    reference point (NF_C_H_Q1_3D_Xi/Eta/Zeta[k],
    include/FE3D/NF_C_H_Q1_3D.h:13-15, resp. NF_C_H_Q2_3D,
    include/FE3D/NF_C_H_Q2_3D.h:13-21), mapped by THexaAffin
    (src/FE/HexaAffin.C:872-875): x = xc0 + xc1 xi + xc2 eta + xc3 zeta
    with xc0 = (x1 + x3 + x4 - x0) / 2, xc1 = (x1 - x0) / 2,
    xc2 = (x3 - x0) / 2, xc3 = (x4 - x0) / 2.
 4. Classes.  With Dirichlet conditions on the whole boundary a DOF is a
    Dirichlet DOF when its nodal point lies on the boundary, else inner.
 5. Numbers.  TFE3DMapper::MapDOF (src/FE/FE3DMapper.C:142-217) joins the
    coinciding DOFs of neighbouring cells so that the root of every class is
    its lowest position in the cell-major array (w = MIN(w0, w1), and every
    link points to a lower position, asserted at :241-257).  The final pass
    of TFESpace3D::ConstructSpace (src/FE/FESpace3D.C:1832-1908) walks the
    positions in order and gives every root the next number of its class
    (inner: count++, Dirichlet: DirichletCounter++), every other position
    its root's number; the offsets (:1931-1995) then put the inner DOFs
    first and the Dirichlet DOFs last.  So a DOF's number is its rank, by
    first position, within its class, inner before Dirichlet.

Usage: python tests/golden/make_hexa_numbering_fixture.py
"""
import json
import os
import re

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hexa_numbering.json")


def read(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read()


def c_array(text, name, typ=float):
    """numbers of the initialiser of the C array `name` (nested braces flattened)"""
    m = re.search(r"\b%s(\s*\[[^\]]*\])+\s*=\s*\{" % re.escape(name), text)
    assert m, name
    i, depth = m.end() - 1, 0
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                body = text[i:j + 1]
                break
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return [typ(v) for v in re.findall(r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", body)]


def c_rows(text, name):
    """rows (inner brace groups) of a two-dimensional C int array"""
    m = re.search(r"\b%s(\s*\[[^\]]*\])+\s*=\s*\{" % re.escape(name), text)
    assert m, name
    body = text[m.end():text.index(";", m.end())]
    return [[int(v) for v in re.findall(r"-?\d+", g)] for g in re.findall(r"\{([^{}]*)\}", body)]


def level0_vertices():
    tests = read("src/Geometry/Tests.C")
    bd = tests[tests.index("int TDomain::initializeDefaultCubeBdry()"):]
    bd = bd[bd.index('"PARAMETERS \\n"'):bd.index("ReadBdParam")]
    nums = [float(v) for line in re.findall(r'bdryStream << "([^"]*)\\n"', bd) for v in line.split()]
    planes = np.array(nums).reshape(6, 3, 3)
    mesh = tests[tests.index("void TDomain::initialize_cube_hexa_mesh()"):]
    mesh = mesh[mesh.index('"DCORVG'):mesh.index('"KVERT')]
    rows = re.findall(r'geoStream << "([^"]*)\\n"', mesh)
    verts = []
    for line in rows:
        comp, t, s = (float(v) for v in line.split())
        P, A, N = planes[int(comp)]
        B = np.cross(A, N)                      # BdPlane.C:127-129
        verts.append(P + t * A + s * B)         # BdPlane.C:49-51
    return np.array(verts)


def refinement():
    rd = read("src/Refinement/RefHexaRegDesc.C")
    hx = read("src/Geometry/Hexahedron.C")
    child_vertex = np.array(c_array(rd, "DatChildVertex", int)).reshape(8, 8)
    new_eq_old = c_array(rd, "DatNewVertexEqOldVertex", int)
    edge_vertex = np.array(c_array(hx, "DatEdgeVertex", int)).reshape(12, 2)
    face_vertex = np.array(c_array(hx, "DatFaceVertex", int)).reshape(6, 4)
    iv_edge = c_array(rd, "DatInteriorVertexOfEdge", int)
    iv_face = c_array(rd, "DatInteriorVertexOfFace", int)
    iv_cell = c_array(rd, "DatInteriorVertexOfCell", int)
    face_pos = np.array(c_array(rd, "DatOldFaceNewVertexPos")).reshape(6, 9, 4)
    int_pos = np.array(c_array(rd, "DatPositionOfIntVert")).reshape(1, 8)
    # weights of the old vertices for each of the 27 new vertices
    W = np.zeros((27, 8))
    for k, v in enumerate(new_eq_old):
        W[v, k] = 1.0
    for e, v in enumerate(iv_edge):
        W[v, edge_vertex[e]] = 0.5
    for f, v in enumerate(iv_face):
        W[v, face_vertex[f]] = face_pos[f][8]      # the face's interior point
    W[iv_cell[0]] = int_pos[0]
    assert np.allclose(W.sum(1), 1.0)
    # check: centroid of the old vertices of the children holding the vertex
    rows = c_rows(rd, "DatVertexChild")
    vlen = c_array(rd, "DatVertexChildLen", int)
    W2 = np.zeros((27, 8))
    assert [len(r) for r in rows] == vlen
    for v, ch in enumerate(rows):
        for c in ch:
            assert child_vertex[c][0] == c       # child c holds old vertex c
            W2[v, c] = 1.0 / len(ch)
    assert np.allclose(W, W2), "refinement tables disagree"
    return child_vertex, W


def mesh_levels(n_levels):
    child_vertex, W = refinement()
    cells = [level0_vertices()]
    levels = [cells]
    for _ in range(n_levels):
        fine = []
        for c in levels[-1]:
            new = W @ c                                  # 27 x 3
            for k in range(8):
                fine.append(new[child_vertex[k]])
        levels.append(fine)
    return levels


def local_points(order):
    name = "NF_C_H_Q%d_3D" % order
    nf = read("include/FE3D/%s.h" % name)
    return np.array([c_array(nf, name + s) for s in ("_Xi", "_Eta", "_Zeta")]).T


def numbering(cells, order):
    ref = local_points(order)
    nodes = []
    for v in cells:
        x0, x1, x3, x4 = v[0], v[1], v[3], v[4]
        xc0 = (x1 + x3 + x4 - x0) * 0.5
        xc1, xc2, xc3 = (x1 - x0) * 0.5, (x3 - x0) * 0.5, (x4 - x0) * 0.5
        for p in ref:
            x = xc0 + xc1 * p[0] + xc2 * p[1] + xc3 * p[2]
            nodes.append(tuple(np.round(x * 2 ** 20).astype(np.int64)))
    first = {}
    for pos, k in enumerate(nodes):
        first.setdefault(k, pos)
    scale = 2 ** 20
    on_bd = {k: any(c == 0 or c == scale for c in k) for k in first}
    roots = sorted(first.items(), key=lambda kv: kv[1])
    inner = [k for k, _ in roots if not on_bd[k]]
    diri = [k for k, _ in roots if on_bd[k]]
    num = {k: i for i, k in enumerate(inner)}
    num.update({k: len(inner) + i for i, k in enumerate(diri)})
    g = np.array([num[k] for k in nodes]).reshape(len(cells), -1)
    return g, len(inner), len(first)


def main():
    levels = mesh_levels(2)
    out = {"_source": __doc__.strip().splitlines()[0] + " (generated by tests/golden/"
                      "make_hexa_numbering_fixture.py from the reference's source tables)"}
    for lev in (1, 2):
        cells = levels[lev]
        out["level%d_vertices" % lev] = np.array(cells).reshape(-1).tolist()
        for order in (1, 2):
            g, n_inner, n_dof = numbering(cells, order)
            out["level%d_Q%d" % (lev, order)] = {"n_dof": n_dof, "n_inner": n_inner,
                                                 "global_numbers": g.reshape(-1).tolist()}
    with open(OUT, "w") as f:
        json.dump(out, f)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
