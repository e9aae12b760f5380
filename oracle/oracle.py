"""ORACLE — test infrastructure only (tests/, smoke(), bench.py's cpu_baseline).

Python/numpy restatement of the reference's set-up path plus ctypes access to
the C restatement of its assembly/solve path (cd_oracle.c):

* mesh(): macro meshes of TDomain::UnitSquare / TwoTriangles /
  initialize_cube_hexa_mesh (src/Geometry/Tests.C:95-170, 449-482, 1815-1856)
  refined with the child-vertex tables of RefQuadRegDesc.C:20-21,
  RefTriRegDesc.C:20-21, RefHexaRegDesc.C:31-35 in It_Finest depth-first order
  (src/Geometry/It_LE.C:17-56).
* number_dofs(): TFESpace3D/2D::ConstructSpace (src/FE/FESpace3D.C:506-2080,
  FESpace2D.C:1284-2000) with the mark-based union-find of TFE3DMapper::MapBound
  / MapDOF (src/FE/FE3DMapper.C:67-220): boundary joints are marked with the
  class counters, joints shared with a later cell are paired (lowest index
  becomes the root, the highest mark wins), inner DOFs get a fresh mark; the
  renumbering pass and the final class offsets follow FESpace3D.C:1786-1990.
  Coinciding local DOFs of neighbouring cells are paired by their nodal
  points (the reference pairs them through its face tables and twists).
* structure(): TStructure(FESpace) (src/Matrix/Structure.C:829-1277).

This module never runs in the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

TRI, QUAD, HEXA = 0, 1, 2
FIRSTMARK = -10

# ------------------------------------------------------------------ meshes
_QUAD_MEAN = [[0], [1], [2], [3], [0, 1], [1, 2], [2, 3], [3, 0], [0, 1, 2, 3]]
_QUAD_CHILD = [[0, 4, 8, 7], [1, 5, 8, 4], [2, 6, 8, 5], [3, 7, 8, 6]]
_TRI_MEAN = [[0], [1], [2], [0, 1], [1, 2], [2, 0]]
_TRI_CHILD = [[0, 3, 5], [1, 4, 3], [2, 5, 4], [3, 4, 5]]
_HEXA_MEAN = [[0], [1], [2], [3], [4], [5], [6], [7], [0, 1], [1, 2], [2, 3], [0, 3],
              [0, 1, 2, 3], [1, 5], [4, 5], [0, 4], [0, 1, 5, 4], [2, 6], [5, 6],
              [1, 2, 6, 5], [3, 7], [6, 7], [2, 3, 7, 6], [4, 7], [0, 3, 7, 4],
              [4, 5, 6, 7], [0, 1, 2, 3, 4, 5, 6, 7]]
_HEXA_CHILD = [[0, 8, 12, 11, 15, 16, 26, 24], [1, 9, 12, 8, 13, 19, 26, 16],
               [2, 10, 12, 9, 17, 22, 26, 19], [3, 11, 12, 10, 20, 24, 26, 22],
               [4, 23, 25, 14, 15, 24, 26, 16], [5, 14, 25, 18, 13, 16, 26, 19],
               [6, 18, 25, 21, 17, 19, 26, 22], [7, 21, 25, 23, 20, 22, 26, 24]]

JOINT_VERTS = {
    QUAD: [[0, 1], [1, 2], [2, 3], [3, 0]],
    TRI: [[0, 1], [1, 2], [2, 0]],
    HEXA: [[0, 1, 2, 3], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7], [4, 7, 6, 5]],
}


def mesh(geometry, n_refine):
    """Return (shape, levels) with levels[k] = dict(verts [nc,nv,dim], parent [nc])."""
    if geometry == 0:
        shape, v0 = QUAD, np.array([[[0, 0], [1, 0], [1, 1], [0, 1]]], float)
        mean, child = _QUAD_MEAN, _QUAD_CHILD
    elif geometry == 1:
        shape, v0 = TRI, np.array([[[0, 0], [1, 0], [1, 1]], [[0, 0], [1, 1], [0, 1]]], float)
        mean, child = _TRI_MEAN, _TRI_CHILD
    else:
        shape = HEXA
        v0 = np.array([[[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]], float)
        mean, child = _HEXA_MEAN, _HEXA_CHILD
    levels = [dict(verts=v0, parent=np.full(len(v0), -1, np.int32))]
    for _ in range(n_refine):
        pv = levels[-1]["verts"]
        new = np.stack([pv[:, m, :].mean(axis=1) for m in mean], axis=1)   # [nc, n_new, dim]
        ch = np.stack([new[:, c, :] for c in child], axis=1)                 # [nc, nch, nv, dim]
        nc, nch = ch.shape[0], ch.shape[1]
        levels.append(dict(verts=ch.reshape(nc * nch, ch.shape[2], ch.shape[3]),
                           parent=np.repeat(np.arange(nc, dtype=np.int32), nch)))
    return shape, levels


def ndof_cell(shape, order):
    if shape == TRI:
        return 3
    return (order + 1) ** (3 if shape == HEXA else 2)


def ref_nodes(shape, order):
    if shape == TRI:
        return np.array([[0, 0], [1, 0], [0, 1]], float)
    n1, dim = order + 1, (3 if shape == HEXA else 2)
    out = []
    for k in range(ndof_cell(shape, order)):
        idx = [k % n1, (k // n1) % n1, k // (n1 * n1)]
        out.append([-1.0 + 2.0 * idx[d] / order for d in range(dim)])
    return np.array(out)


def node_phys(shape, verts, xi):
    """nodal point of a parallelogram/parallelepiped/triangle cell (vertex-weighted map)"""
    if shape == TRI:
        return verts[0] + (verts[1] - verts[0]) * xi[0] + (verts[2] - verts[0]) * xi[1]
    if shape == QUAD:
        s = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
        return sum(0.25 * (1 + a * xi[0]) * (1 + b * xi[1]) * verts[k] for k, (a, b) in enumerate(s))
    s = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    return sum(0.125 * (1 + a * xi[0]) * (1 + b * xi[1]) * (1 + c * xi[2]) * verts[k]
               for k, (a, b, c) in enumerate(s))


def boundary_component(shape, fv):
    """boundary component of a joint with vertex coordinates fv, or -1"""
    if shape in (QUAD, TRI):
        a, b = fv
        if a[1] == 0 and b[1] == 0:
            return 0
        if a[0] == 1 and b[0] == 1:
            return 1
        if a[1] == 1 and b[1] == 1:
            return 2
        if a[0] == 0 and b[0] == 0:
            return 3
        return -1
    for f, (ax, val) in enumerate([(2, 0), (1, 0), (0, 1), (1, 1), (0, 0), (2, 1)]):
        if all(p[ax] == val for p in fv):
            return f
    return -1


def bparam(comp, x, y):
    return [x, y, 1 - x, 1 - y][comp]


def joint_dofs(shape, order, j):
    xi = ref_nodes(shape, order)
    out = []
    for k, p in enumerate(xi):
        if shape == TRI:
            on = [p[1] == 0, p[0] + p[1] == 1, p[0] == 0][j]
        elif shape == QUAD:
            on = p[[1, 0, 1, 0][j]] == [-1, 1, 1, -1][j]
        else:
            on = p[[2, 1, 0, 1, 0, 2][j]] == [-1, -1, 1, 1, -1, 1][j]
        if on:
            out.append(k)
    return out


def number_dofs(shape, verts, order, cond2d=None, cond3d=None):
    """ConstructSpace: returns (global_numbers [nc, nd], info dict, bd_comp [nc, nj])."""
    nc, nd = len(verts), ndof_cell(shape, order)
    nj = len(JOINT_VERTS[shape])
    xi = ref_nodes(shape, order)
    key = np.zeros((nc, nd), dtype=np.int64)
    coords = {}
    keymap = {}
    for c in range(nc):
        for k in range(nd):
            x = tuple(np.round(node_phys(shape, verts[c], xi[k]), 12))
            if x not in keymap:
                keymap[x] = len(keymap)
                coords[keymap[x]] = x
            key[c, k] = keymap[x]
    # joints: boundary component and neighbour (cell, local joint)
    bd = np.full((nc, nj), -1, np.int32)
    face_owner = {}
    neigh = {}
    for c in range(nc):
        for j, fvi in enumerate(JOINT_VERTS[shape]):
            fv = [tuple(verts[c][v]) for v in fvi]
            comp = boundary_component(shape, fv)
            if comp >= 0:
                bd[c, j] = comp
                continue
            fk = tuple(sorted(fv))
            if fk in face_owner:
                c2, j2 = face_owner[fk]
                neigh[(c, j)] = (c2, j2)
                neigh[(c2, j2)] = (c, j)
            else:
                face_owner[fk] = (c, j)
    jd = [joint_dofs(shape, order, j) for j in range(nj)]

    def cond_of(c, j):
        comp = bd[c, j]
        fvi = JOINT_VERTS[shape][j]
        if shape == HEXA:
            X = np.mean([verts[c][v] for v in fvi], axis=0)
            return cond3d(X[0], X[1], X[2])
        a, b = verts[c][fvi[0]], verts[c][fvi[1]]
        t0, t1 = bparam(comp, a[0], a[1]), bparam(comp, b[0], b[1])
        eps = 1e-6
        c0, c1 = (cond2d(comp, t0 + eps), cond2d(comp, t1 - eps)) if t0 < t1 else \
                 (cond2d(comp, t0 - eps), cond2d(comp, t1 + eps))
        assert c0 == c1
        return c0

    # upper bounds of the boundary classes (first pass of ConstructSpace)
    NT = 9
    dir_ub, bnd_ub = 0, [0] * NT
    for c in range(nc):
        for j in range(nj):
            if bd[c, j] >= 0:
                cond = cond_of(c, j)
                if cond == 0:
                    dir_ub += len(jd[0])
                else:
                    bnd_ub[cond - 1] += len(jd[0])
    G = np.full(nc * nd, -1, dtype=np.int64)
    sum_loc = nc * nd
    counters = {"dir": FIRSTMARK + 1}
    l = -dir_ub + FIRSTMARK + 1
    bound_counter = []
    for t in range(NT):
        bound_counter.append(l)
        l -= bnd_ub[t]
    counters["inner"] = l

    def root(w):
        while G[w] > -1:
            w = G[w]
        return w

    def map_bound(c, dofs, which):
        for k in dofs:
            dof = c * nd + k
            w = root(dof)
            cur = counters["dir"] if which == "dir" else bound_counter[which]
            if G[w] == -1:
                cur -= 1
                G[dof] = cur
            elif G[w] < cur:
                cur -= 1
                G[w] = cur
                if dof != w:
                    G[dof] = w
            if which == "dir":
                counters["dir"] = cur
            else:
                bound_counter[which] = cur

    def map_dof(d0, d1):
        w0, w1 = root(d0), root(d1)
        if G[d0] == -1 and G[d1] == -1:
            counters["inner"] -= 1
            if d0 < d1:
                G[d0] = counters["inner"]; G[d1] = d0
            else:
                G[d1] = counters["inner"]; G[d0] = d1
        elif G[d0] == -1:
            if w1 > d0:
                G[d0] = G[w1]; G[d1] = d0; G[w1] = d0
            else:
                G[d0] = w1
        elif G[d1] == -1:
            if w0 > d1:
                G[d1] = G[w0]; G[d0] = d1; G[w0] = d1
            else:
                G[d1] = w0
        else:
            if G[w0] != G[w1]:
                e = max(G[w0], G[w1])
                w = min(w0, w1)
                G[w0] = w; G[w1] = w; G[w] = e

    for c in range(nc):
        for j in range(nj):
            if bd[c, j] >= 0:
                cond = cond_of(c, j)
                map_bound(c, jd[j], "dir" if cond == 0 else cond - 1)
            else:
                c2, j2 = neigh[(c, j)]
                if c2 > c:
                    for k in jd[j]:
                        for k2 in jd[j2]:
                            if key[c, k] == key[c2, k2]:
                                map_dof(c * nd + k, c2 * nd + k2)
        # inner dofs: nodes on no joint
        on_joint = set(k for j in range(nj) for k in jd[j])
        for k in range(nd):
            if k not in on_joint:
                counters["inner"] -= 1
                G[c * nd + k] = counters["inner"]
    # renumbering pass (FESpace3D.C:1804-1905)
    bound_mark = []
    l = FIRSTMARK - dir_ub
    for t in range(NT):
        bound_mark.append(l)
        l -= bnd_ub[t]
    inner_mark = l
    dcount = 0
    bcount = []
    l = dir_ub
    for t in range(NT):
        bcount.append(l)
        l += bnd_ub[t]
    count = l
    n_dir, n_inner, n_bnd = 0, 0, [0] * NT
    for i in range(sum_loc):
        v = G[i]
        if v < -1:
            if v <= inner_mark:
                G[i] = count; count += 1; n_inner += 1
                continue
            m = NT - 1
            matched = False
            while m >= 0:
                if v <= bound_mark[m]:
                    G[i] = bcount[m]; bcount[m] += 1; n_bnd[m] += 1
                    matched = True
                    break
                m -= 1
            if not matched:
                G[i] = dcount; dcount += 1; n_dir += 1
        else:
            G[i] = G[v]
    # class offsets (FESpace3D.C:1923-1975)
    m_ub = sum(bnd_ub)
    dir_off = n_inner + sum(n_bnd)
    inner_off = -dir_ub - m_ub
    bound_off, l, m = [], n_inner, dir_ub
    for t in range(NT):
        bound_off.append(l - m)
        l += n_bnd[t]
        m += bnd_ub[t]
    bmark, l = [], dir_ub
    for t in range(NT):
        l += bnd_ub[t]
        bmark.append(l)
    out = np.empty(sum_loc, dtype=np.int64)
    for i in range(sum_loc):
        n = G[i]
        if n < dir_ub:
            out[i] = n + dir_off
        elif n < bmark[NT - 1]:
            for t in range(NT):
                if n < bmark[t]:
                    out[i] = n + bound_off[t]
                    break
        else:
            out[i] = n + inner_off
    active = n_inner + sum(n_bnd)
    info = dict(n_dof=active + n_dir, active_bound=active, n_inner=n_inner, n_bound=n_bnd,
                n_dirichlet=n_dir)
    return out.reshape(nc, nd).astype(np.int32), info, bd


def structure(gnum, n_dof, active):
    """TStructure(FESpace): active rows = sorted union of cell DOFs; Dirichlet rows = diagonal."""
    rows = [set() for _ in range(n_dof)]
    for cell in gnum:
        for a in cell:
            if a < active:
                rows[a].update(int(b) for b in cell)
    rp = [0]
    col = []
    for r in range(n_dof):
        cols = sorted(rows[r]) if r < active else [r]
        col.extend(cols)
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32)


# ------------------------------------------------------------- C oracle
class Space(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int), ("order", ctypes.c_int), ("nv", ctypes.c_int),
                ("n_cells", ctypes.c_int), ("ndof_cell", ctypes.c_int), ("n_dof", ctypes.c_int),
                ("active", ctypes.c_int), ("verts", ctypes.c_void_p), ("gnum", ctypes.c_void_p),
                ("rp", ctypes.c_void_p), ("col", ctypes.c_void_p), ("bd_comp", ctypes.c_void_p)]


class CD(ctypes.Structure):
    _fields_ = [("disc", ctypes.c_int), ("example", ctypes.c_int), ("hexa_rule", ctypes.c_int),
                ("bc_sine_laplace", ctypes.c_int), ("eps", ctypes.c_double),
                ("b", ctypes.c_double * 3), ("c", ctypes.c_double), ("f0", ctypes.c_double),
                ("delta0", ctypes.c_double), ("poly_degree", ctypes.c_double),
                ("time", ctypes.c_double), ("theta1", ctypes.c_double), ("tau", ctypes.c_double),
                ("table", ctypes.c_void_p)]


class SolverParams(ctypes.Structure):
    _fields_ = [("iterative_solver_type", ctypes.c_int), ("preconditioner", ctypes.c_int),
                ("max_n_iterations", ctypes.c_int), ("min_n_iterations", ctypes.c_int),
                ("residual_tolerance", ctypes.c_double), ("residual_reduction", ctypes.c_double),
                ("divergence_factor", ctypes.c_double), ("gmres_restart", ctypes.c_int),
                ("damping", ctypes.c_double), ("mg_cycle", ctypes.c_int),
                ("mg_smoother_coarse", ctypes.c_int), ("mg_correction_damp", ctypes.c_double),
                ("mg_n_pre_smooth", ctypes.c_int), ("mg_n_post_smooth", ctypes.c_int),
                ("mg_coarse_residual", ctypes.c_double), ("mg_coarse_max_n_iterations", ctypes.c_int),
                ("jacobi_damp", ctypes.c_double)]


class Level(ctypes.Structure):
    _fields_ = [("s", Space), ("val", ctypes.c_void_p), ("parent", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(LIB)
        L.or_assemble_cd.argtypes = [ctypes.POINTER(Space), ctypes.POINTER(CD), ctypes.c_void_p, ctypes.c_void_p]
        L.or_errors.argtypes = [ctypes.POINTER(Space), ctypes.POINTER(CD), ctypes.c_void_p, ctypes.c_void_p]
        L.or_quad_points.argtypes = [ctypes.POINTER(Space), ctypes.c_int, ctypes.c_void_p]
        L.or_quad_points.restype = ctypes.c_int
        L.or_assemble_tcd.argtypes = [ctypes.POINTER(Space), ctypes.POINTER(CD), ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]
        L.or_interpolate.argtypes = [ctypes.POINTER(Space), ctypes.POINTER(CD), ctypes.c_void_p]
        L.or_multiply_active.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
        L.or_assemble_nse.argtypes = [ctypes.POINTER(Space), ctypes.POINTER(Space), ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.or_assemble_nse2.argtypes = [ctypes.POINTER(Space), ctypes.POINTER(Space), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double] + [ctypes.c_void_p] * 8
        L.or_apply_scaled_add.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
        L.or_prolongate.argtypes = [ctypes.POINTER(Space), ctypes.POINTER(Space), ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p]
        L.or_restrict.argtypes = L.or_prolongate.argtypes
        L.or_solve.argtypes = [ctypes.POINTER(SolverParams), ctypes.c_int, ctypes.POINTER(Level),
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                               ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.or_mg_apply.argtypes = [ctypes.POINTER(SolverParams), ctypes.c_int, ctypes.POINTER(Level),
                                  ctypes.c_void_p, ctypes.c_void_p]
        L.or_basis.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p]
        L.or_quad_rule.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleSpace:
    """Arrays of one FE space kept alive for the C structs."""

    def __init__(self, shape, order, verts, gnum, n_dof, active, rp, col, bd, parent=None):
        self.shape, self.order = shape, order
        self.verts = np.ascontiguousarray(verts, dtype=np.float64)
        self.gnum = np.ascontiguousarray(gnum, dtype=np.int32)
        self.rp = np.ascontiguousarray(rp, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.bd = np.ascontiguousarray(bd, dtype=np.int32)
        self.parent = None if parent is None else np.ascontiguousarray(parent, dtype=np.int32)
        self.n_dof, self.active = int(n_dof), int(active)
        self.n_cells, self.ndof_cell = self.gnum.shape
        self.nnz = len(self.col)

    def cstruct(self):
        return Space(self.shape, self.order, self.verts.shape[1], self.n_cells, self.ndof_cell,
                     self.n_dof, self.active, _p(self.verts), _p(self.gnum), _p(self.rp),
                     _p(self.col), _p(self.bd))


def build_space(geometry, n_refine, level, order, cond2d=None, cond3d=None):
    shape, levels = mesh(geometry, n_refine)
    v = levels[level]["verts"]
    if cond2d is None:
        cond2d = lambda comp, t: 0   # noqa: E731
    if cond3d is None:
        cond3d = lambda x, y, z: 0   # noqa: E731
    g, info, bd = number_dofs(shape, v, order, cond2d, cond3d)
    rp, col = structure(g, info["n_dof"], info["active_bound"])
    return OracleSpace(shape, order, v, g, info["n_dof"], info["active_bound"], rp, col, bd,
                       levels[level]["parent"]), info


def cd_params(example=0, disc=0, eps=1.0, b=(0, 0, 0), c=0.0, f0=0.0, hexa_rule=0,
              sine_laplace_bc=False, delta0=1.0, poly_degree=1.0, time=0.0, theta1=1.0, tau=0.0,
              table=None):
    """`table`: optional [n_cells, n_quad, 8] per-point coefficients (eps, b1,
    b2, b3, c, f, -, -) at the points of quad_points(); kept alive on the
    returned struct."""
    p = CD()
    if table is not None:
        p._table = np.ascontiguousarray(table, dtype=np.float64)
        p.table = p._table.ctypes.data
    p.disc, p.example, p.hexa_rule, p.bc_sine_laplace = disc, example, hexa_rule, int(sine_laplace_bc)
    p.eps, p.c, p.f0, p.delta0, p.poly_degree = eps, c, f0, delta0, poly_degree
    p.time, p.theta1, p.tau = time, theta1, tau
    for k in range(3):
        p.b[k] = b[k] if k < len(b) else 0.0
    return p


def quad_points(space, hexa_rule=0):
    """physical quadrature points [n_cells, n_quad, 3] of the assembly loop"""
    out = np.zeros((space.n_cells, 64, 3))
    s = space.cstruct()
    nq = lib().or_quad_points(ctypes.byref(s), hexa_rule, _p(out))
    return np.ascontiguousarray(out.reshape(-1)[: space.n_cells * nq * 3].reshape(space.n_cells, nq, 3))


def assemble(space, params):
    val = np.zeros(space.nnz)
    rhs = np.zeros(space.n_dof)
    s = space.cstruct()
    lib().or_assemble_cd(ctypes.byref(s), ctypes.byref(params), _p(val), _p(rhs))
    return val, rhs


def assemble_tcd(space, params, stiffness=True, mass=True, rhs=True):
    """TCD stiffness, mass and rhs (each None when not requested)."""
    a = np.zeros(space.nnz) if stiffness else None
    m = np.zeros(space.nnz) if mass else None
    f = np.zeros(space.n_dof) if rhs else None
    s = space.cstruct()
    lib().or_assemble_tcd(ctypes.byref(s), ctypes.byref(params), None if a is None else _p(a),
                          None if m is None else _p(m), None if f is None else _p(f))
    return a, m, f


def interpolate(space, params):
    out = np.zeros(space.n_dof)
    s = space.cstruct()
    lib().or_interpolate(ctypes.byref(s), ctypes.byref(params), _p(out))
    return out


def multiply_active(active, rp, col, val, x, y, a):
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    assert y.dtype == np.float64 and y.flags["C_CONTIGUOUS"]
    lib().or_multiply_active(active, _p(rp), _p(col), _p(val), _p(x), _p(y), a)
    return y


BACKWARD_EULER, CRANK_NICOLSON, BDF2 = 0, 1, 2


def tcd_time_integration(space, example, disc, scheme, tau, n_steps, eps=1.0, b=(0.0, 0.0), c=0.0,
                         t0=0.0, solver=None, on_step=None):
    """Restatement of TimeConvectionDiffusion<2> (src/System/TimeConvectionDiffusion.C:
    294-386 assemble/solve, 222-290 assemble_initial_time) driven like
    test/tcd2d_test.c++:103-132, with the TimeDiscretization steps
    prepare_rhs_from_time_disc / prepare_system_matrix
    (src/TimeDiscRoutines/TimeDiscretizations.C:106-210).  THETA1 follows
    TimeDiscretization::set_time_disc_parameters (:341-381).  `solver(S, rhs, x0)`
    returns the solution of the system (default: FGMRES + Jacobi to 1e-14).
    Returns the (L2, H1) errors after the initial interpolation and after every step."""
    ab, rp, col = space.active, space.rp, space.col
    na = int(rp[ab])
    supg = disc == 1

    def theta(step):
        if scheme == CRANK_NICOLSON:
            return 0.5
        if scheme == BDF2 and step >= 2:
            return 2.0 / 3.0
        return 1.0

    def params(t, step):
        return cd_params(example=example, disc=disc, eps=eps, b=tuple(b) + (0.0,), c=c, time=t,
                         theta1=theta(step), tau=tau)

    if solver is None:
        sp = solver_params(iterative_solver_type=0, preconditioner=1, gmres_restart=50,
                           max_n_iterations=2000, residual_tolerance=1e-14, residual_reduction=0.0)

        def solver(S, rhs, x0):
            x, it, _, _ = solve(sp, [space], [S], rhs, x0)
            return x

    t = t0
    p = params(t, 0)
    u = interpolate(space, p)
    A, M, f = assemble_tcd(space, p)
    old_rhs = f.copy()
    u_m1, u_m2 = u.copy(), u.copy()
    errs = [errors(space, p, u)[:2]]
    for step in range(1, n_steps + 1):
        t += tau
        p = params(t, step)
        if supg:
            A, M, f = assemble_tcd(space, p)
        else:
            A, _, f = assemble_tcd(space, p, mass=False)
        rhs_td = f.copy()
        bdf2_full = scheme == BDF2 and step >= 2
        if scheme == BACKWARD_EULER or (scheme == BDF2 and not bdf2_full):
            rhs_td[:ab] *= tau
            multiply_active(ab, rp, col, M, u_m1, rhs_td, 1.0)
            factor = tau
        elif scheme == CRANK_NICOLSON:
            rhs_td[:ab] *= 0.5 * tau
            rhs_td[:ab] += 0.5 * tau * old_rhs[:ab]
            multiply_active(ab, rp, col, M, u_m1, rhs_td, 1.0)
            multiply_active(ab, rp, col, A, u_m1, rhs_td, -0.5 * tau)
            factor = 0.5 * tau
        else:
            rhs_td[:ab] *= (2.0 / 3.0) * tau
            multiply_active(ab, rp, col, M, u_m1, rhs_td, 4.0 / 3.0)
            multiply_active(ab, rp, col, M, u_m2, rhs_td, -1.0 / 3.0)
            factor = (2.0 / 3.0) * tau
        old_rhs = f.copy()
        rhs_td[ab:] = f[ab:]
        S = A.copy()
        S[:na] *= factor
        S[:na] += M[:na]
        u[ab:] = f[ab:]
        u = solver(S, rhs_td, u)
        u_m2, u_m1 = u_m1, u.copy()
        errs.append(errors(space, p, u)[:2])
        if on_step is not None:
            on_step(step, t, u)
    return np.array(errs), u


def errors(space, params, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(4)
    s = space.cstruct()
    lib().or_errors(ctypes.byref(s), ctypes.byref(params), _p(u), _p(out))
    return out


def apply_scaled_add(n, active, rp, col, val, x, y, a):
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    assert y.dtype == np.float64 and y.flags["C_CONTIGUOUS"]
    lib().or_apply_scaled_add(n, active, _p(rp), _p(col), _p(val), _p(x), _p(y), a)
    return y


def prolongate(coarse, fine, u):
    u = np.ascontiguousarray(u, np.float64)
    out = np.zeros(fine.n_dof)
    cs, fs = coarse.cstruct(), fine.cstruct()
    lib().or_prolongate(ctypes.byref(cs), ctypes.byref(fs), _p(fine.parent), _p(u), _p(out))
    return out


def restrict(coarse, fine, d):
    d = np.ascontiguousarray(d, np.float64)
    out = np.zeros(coarse.n_dof)
    cs, fs = coarse.cstruct(), fine.cstruct()
    lib().or_restrict(ctypes.byref(cs), ctypes.byref(fs), _p(fine.parent), _p(out), _p(d))
    return out


def solver_params(iterative_solver_type=0, preconditioner=0, max_n_iterations=100,
                  min_n_iterations=0, residual_tolerance=1e-8, residual_reduction=0.0,
                  divergence_factor=1e10, gmres_restart=20, damping=1.0, mg_cycle=0,
                  mg_smoother_coarse=1, mg_correction_damp=1.0, mg_n_pre_smooth=1,
                  mg_n_post_smooth=1, mg_coarse_residual=0.1, mg_coarse_max_n_iterations=10,
                  jacobi_damp=1.0):
    return SolverParams(iterative_solver_type, preconditioner, max_n_iterations, min_n_iterations,
                        residual_tolerance, residual_reduction, divergence_factor, gmres_restart,
                        damping, mg_cycle, mg_smoother_coarse, mg_correction_damp, mg_n_pre_smooth,
                        mg_n_post_smooth, mg_coarse_residual, mg_coarse_max_n_iterations, jacobi_damp)


def _levels(spaces, vals):
    keep = []
    arr = (Level * len(spaces))()
    for k, (s, v) in enumerate(zip(spaces, vals)):
        v = np.ascontiguousarray(v, np.float64)
        keep.append(v)
        arr[k].s = s.cstruct()
        arr[k].val = _p(v)
        arr[k].parent = _p(s.parent) if (k > 0 and s.parent is not None) else None
    return arr, keep


def solve(params, spaces, vals, b, x0):
    """spaces/vals of the multigrid levels, coarsest first (or a single level)."""
    arr, keep = _levels(spaces, vals)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, dtype=np.float64, copy=True)
    it = ctypes.c_int()
    hist = np.zeros(100000)
    nh = ctypes.c_int()
    st = lib().or_solve(ctypes.byref(params), len(spaces), arr, _p(b), _p(x), ctypes.byref(it),
                        _p(hist), len(hist), ctypes.byref(nh))
    return x, it.value, hist[: min(nh.value, len(hist))], st


def mg_apply(params, spaces, vals, v):
    arr, keep = _levels(spaces, vals)
    v = np.ascontiguousarray(v, np.float64)
    z = np.zeros_like(v)
    lib().or_mg_apply(ctypes.byref(params), len(spaces), arr, _p(v), _p(z))
    return z


def csr_general(n, active, rp, col):
    """OracleSpace-like holder for a plain CSR matrix (no mesh)."""
    return OracleSpace(HEXA, 1, np.zeros((1, 8, 3)), np.zeros((1, 8), np.int32), n, active, rp, col,
                       np.full((1, 6), -1, np.int32))


# --------------------------------------------------- Navier-Stokes (NSTYPE 1)
def rect_structure(gtest, gansatz, n_test):
    """TStructure(testspace, ansatzspace) (src/Matrix/Structure.C:1709-1924):
    row t = sorted union of the ansatz DOFs of every cell containing t."""
    rows = [set() for _ in range(n_test)]
    for ct, ca in zip(gtest, gansatz):
        for t in ct:
            rows[int(t)].update(int(a) for a in ca)
    rp, col = [0], []
    for r in range(n_test):
        col.extend(sorted(rows[r]))
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32)


def front_face_interior(x, y, z):
    """test_u_0_p_0.h on_front_face_interior: the face x = 1 without its boundary."""
    tol = 1e-10
    return abs(1 - x) < tol and not abs(y - 1) < tol and not abs(y) < tol and \
        not abs(z - 1) < tol and not abs(z) < tol


class NseOracle:
    """Velocity Q2 and pressure Q1 (continuous, Neumann) or discontinuous P1
    (D_P1_3D_H_A: 4 DOFs per cell numbered 4 cell + local) spaces of one
    hexahedral level with the patterns of A and B.  example "u2_p1"
    (test_u_2_p_1.h): Dirichlet everywhere; "u0_p0" (test_u_0_p_0.h):
    Neumann on the interior of the face x = 1, judged at the centre of each
    boundary face as ConstructSpace does."""

    def __init__(self, n_refine, level, pressure="q1", example="u2_p1"):
        shape, levels = mesh(2, n_refine)
        v = levels[level]["verts"]
        self.pressure, self.example = pressure, example
        if example == "u0_p0":
            vcond = lambda x, y, z: 1 if front_face_interior(x, y, z) else 0   # noqa: E731
        else:
            vcond = lambda x, y, z: 0   # noqa: E731
        gv, iv, bdv = number_dofs(shape, v, 2, None, vcond)
        rp, col = structure(gv, iv["n_dof"], iv["active_bound"])
        self.vs = OracleSpace(shape, 2, v, gv, iv["n_dof"], iv["active_bound"], rp, col, bdv,
                              levels[level]["parent"])
        if pressure == "p1disc":
            nc = len(v)
            gp = np.arange(4 * nc, dtype=np.int32).reshape(nc, 4)
            n_p = 4 * nc
            self.ps = OracleSpace(shape, 1, v, gp, n_p, n_p, np.zeros(n_p + 1, np.int32), np.zeros(0, np.int32),
                                  np.zeros((nc, 6), np.int32), levels[level]["parent"])
        else:
            gp, ip, bdp = number_dofs(shape, v, 1, None, lambda x, y, z: 1)
            n_p = ip["n_dof"]
            prp, pcol = structure(gp, n_p, n_p)
            self.ps = OracleSpace(shape, 1, v, gp, n_p, n_p, prp, pcol, bdp, levels[level]["parent"])
        self.brp, self.bcol = rect_structure(gp, gv, n_p)
        self.n_u, self.n_u_active, self.n_p = iv["n_dof"], iv["active_bound"], n_p
        self.n = 3 * self.n_u + self.n_p
        # BlockFEMatrix.C:1869-1920: enclosed flow = no Neumann / Robin velocity DOF
        self.enclosed = iv["n_bound"][0] == 0 and iv["n_bound"][1] == 0

    def assemble(self, nu, w=None):
        """NSTYPE-1 blocks A, (B1, B2, B3) and rhs of the example (or_assemble_nse2)."""
        A = np.zeros(self.vs.nnz)
        B = [np.zeros(len(self.bcol)) for _ in range(3)]
        rhs = np.zeros(self.n)
        vs, ps = self.vs.cstruct(), self.ps.cstruct()
        wv = None if w is None else _p(np.ascontiguousarray(w, np.float64))
        keep = w
        lib().or_assemble_nse2(ctypes.byref(vs), ctypes.byref(ps), self.ps.ndof_cell,
                               1 if self.example == "u0_p0" else 0, nu, wv, _p(A), _p(self.brp), _p(self.bcol),
                               _p(B[0]), _p(B[1]), _p(B[2]), _p(rhs))
        del keep
        return A, B, rhs

    def pressure_mean(self, p):
        """Mean value of a pressure function on the (axis-parallel hexahedral)
        cells: Q1 the vertex mean per brick, discontinuous P1 its constant
        coefficient (xi, eta, zeta integrate to 0)."""
        if self.pressure == "p1disc":
            vol = np.prod(self.ps.verts.max(1) - self.ps.verts.min(1), axis=1)
            return float(np.sum(vol * p[self.ps.gnum[:, 0]]) / np.sum(vol))
        return q1_mean(self.ps, p)

    def apply(self, A, B, x):
        """BlockFEMatrix::apply (src/Matrix/BlockFEMatrix.C:855-924) for NSTYPE 1:
        multiplyActive of A, multiplyTransposedActive of Bk, identity on
        Dirichlet velocity rows."""
        import scipy.sparse as sp
        nu, na, npr = self.n_u, self.n_u_active, self.n_p
        Am = sp.csr_matrix((A, self.vs.col, self.vs.rp), shape=(nu, nu))
        Bm = [sp.csr_matrix((b, self.bcol, self.brp), shape=(npr, nu)) for b in B]
        y = np.zeros(self.n)
        xp = x[3 * nu:]
        for k in range(3):
            xk = x[k * nu:(k + 1) * nu]
            yk = np.zeros(nu)
            yk[:na] = (Am @ xk)[:na] + (Bm[k].T @ xp)[:na]
            yk[na:] = xk[na:]
            y[k * nu:(k + 1) * nu] = yk
        y[3 * nu:] = sum(Bm[k] @ x[k * nu:(k + 1) * nu] for k in range(3))
        return y

    def combined(self, A, B):
        """BlockFEMatrix::get_combined_matrix (src/Matrix/BlockFEMatrix.C:1168-1345):
        unit Dirichlet rows, first pressure row a unit row."""
        import scipy.sparse as sp
        nu, na, npr = self.n_u, self.n_u_active, self.n_p
        Am = sp.csr_matrix((A, self.vs.col, self.vs.rp), shape=(nu, nu)).tolil()
        Bm = [sp.csr_matrix((b, self.bcol, self.brp), shape=(npr, nu)) for b in B]
        Z = sp.csr_matrix((nu, nu))
        blocks = [[Am if i == j else Z for j in range(3)] + [Bm[i].T] for i in range(3)]
        blocks.append([Bm[0], Bm[1], Bm[2], sp.csr_matrix((npr, npr))])
        K = sp.bmat(blocks, format="lil")
        for k in range(3):
            for r in range(na, nu):
                row = k * nu + r
                K.rows[row] = [row]
                K.data[row] = [1.0]
        if self.enclosed:
            row = 3 * nu
            K.rows[row] = [row]
            K.data[row] = [1.0]
        return K.tocsr()


def p1disc_nf_points():
    """The 8 points of the discontinuous P1 nodal functionals
    (NF_D_H_P1_3D.h): (+-1/sqrt 3)^3, xi fastest."""
    a = 0.5773502691896257645091489
    return np.array([[a if i & 1 else -a, a if i & 2 else -a, a if i & 4 else -a] for i in range(8)])


def p1disc_nf(values8):
    """NF_D_H_P1_3D_EvalAll: F0 = mean of the 8 values, F1..F3 = 0.375 sum of
    value x xi / eta / zeta of the point.  values8 [..., 8] -> [..., 4]."""
    pts = p1disc_nf_points()
    v = np.asarray(values8)
    return np.concatenate([v.mean(-1, keepdims=True), 0.375 * (v @ pts)], -1)


def _affine(verts):
    x0 = node_phys(HEXA, verts, np.zeros(3))
    J = np.stack([node_phys(HEXA, verts, e) - x0 for e in np.eye(3)], -1)
    return x0, J


def p1disc_prolongation(cs, fs):
    """Prolongation of the discontinuous P1 pressure (GridTransfer::Prolongate
    with the matrices of TFEDatabase3D::GetProlongationMatrix3D): the row of
    fine DOF k of child fc holds the child's nodal functional k applied to
    the parent's basis functions 1, xi, eta, zeta; each fine DOF lies in one
    cell, so nothing is averaged.  Returns a scipy CSR matrix [n_f, n_c]."""
    import scipy.sparse as sp
    pts = p1disc_nf_points()
    rows, cols, vals = [], [], []
    for fc in range(fs.n_cells):
        pc = int(fs.parent[fc])
        f0, fJ = _affine(fs.verts[fc])
        c0, cJ = _affine(cs.verts[pc])
        x = f0 + pts @ fJ.T
        xi = np.linalg.solve(cJ, (x - c0).T).T                       # [8, 3] parent reference coords
        basis = np.concatenate([np.ones((8, 1)), xi], 1)             # [8, 4] parent basis at the points
        F = p1disc_nf(basis.T)                                       # [4 coarse, 4 fine functionals]
        for k in range(4):
            for l in range(4):
                if abs(F[l, k]) > 1e-14:
                    rows.append(int(fs.gnum[fc, k]))
                    cols.append(int(cs.gnum[pc, l]))
                    vals.append(F[l, k])
    return sp.csr_matrix((vals, (rows, cols)), shape=(fs.n_dof, cs.n_dof))


def pressure_prolongate(cns, fns, u):
    if cns.pressure == "p1disc":
        return p1disc_prolongation(cns.ps, fns.ps) @ u
    return prolongate(cns.ps, fns.ps, u)


def pressure_restrict(cns, fns, d):
    if cns.pressure == "p1disc":
        return p1disc_prolongation(cns.ps, fns.ps).T @ d
    return restrict(cns.ps, fns.ps, d)


def nse_exact(x):
    """test_u_2_p_1.h: u = (y^2, xz, -y^2), p = x + y + z - 1.5 at points x [n, 3]."""
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    return np.stack([Y * Y, X * Z, -Y * Y]), X + Y + Z - 1.5


def nodal_coords(space):
    """Nodal point of every DOF of an OracleSpace."""
    xi = ref_nodes(space.shape, space.order)
    out = np.zeros((space.n_dof, 3))
    for c in range(space.n_cells):
        for k in range(space.ndof_cell):
            out[space.gnum[c, k]] = node_phys(space.shape, space.verts[c], xi[k])
    return out


def nse_exact_u0p0(x):
    """test_u_0_p_0.h: u = (1, 0, 0), p = 0 at points x [n, 3]."""
    z = np.zeros(len(x))
    return np.stack([z + 1.0, z, z]), z


def nse_interpolate(ns):
    """The example's exact solution in the spaces: velocity at the Q2 nodal
    points, pressure at the Q1 nodal points or through the discontinuous P1
    nodal functionals of every cell."""
    exact = nse_exact_u0p0 if getattr(ns, "example", "u2_p1") == "u0_p0" else nse_exact
    u, _ = exact(nodal_coords(ns.vs))
    if getattr(ns, "pressure", "q1") == "p1disc":
        pts = p1disc_nf_points()
        p = np.zeros(ns.n_p)
        for c in range(ns.ps.n_cells):
            x0, J = _affine(ns.ps.verts[c])
            _, pv = exact(x0 + pts @ J.T)
            p[ns.ps.gnum[c]] = p1disc_nf(pv)
    else:
        _, p = exact(nodal_coords(ns.ps))
    return np.concatenate([u[0], u[1], u[2], p])


# ------------------------------------------- cell Vanka and saddle-point multigrid
class NseLevelOracle:
    """Multigrid level of NavierStokes<3> (NSTYPE 1): system, combined matrix
    and the cell-Vanka patches of VankaSmoother CELL_JACOBI
    (src/Multigrid/VankaSmoother.C:44-300)."""

    def __init__(self, ns, A, B):
        import scipy.sparse  # noqa: F401
        self.ns, self.A, self.B = ns, A, B
        self.K = ns.combined(A, B)
        nu = ns.n_u
        maps = []
        for c in range(ns.vs.n_cells):
            v = np.sort(ns.vs.gnum[c])
            p = np.sort(ns.ps.gnum[c])
            maps.append(np.concatenate([v, v + nu, v + 2 * nu, p + 3 * nu]))
        self.maps = maps
        self.count = np.zeros(ns.n)
        for m in maps:
            self.count[m] += 1
        self.local = [self.K[m, :][:, m].toarray() for m in maps]
        # colour-major cell order of the multiplicative variant: colour =
        # cell lattice index mod 3 in each direction (27 colours), so that the
        # cells of one colour neither share DOFs nor read each other's DOFs
        ctr = ns.vs.verts.mean(1)
        h = ns.vs.verts[0].max(0) - ns.vs.verts[0].min(0)
        ijk = np.floor(ctr / h).astype(int)
        colour = (ijk[:, 0] % 3) + 3 * (ijk[:, 1] % 3) + 9 * (ijk[:, 2] % 3)
        self.colour_order = [int(c) for c in np.lexsort((np.arange(len(colour)), colour))]

    def vanka_mult(self, rhs, x, damp):
        """One CELL sweep (multiplicative, VankaSmoother.C:350-420 with another
        cell order): per cell, in colour-major order, the local rows of the
        current defect, the local solve, and x += damp * local solution."""
        x = x.copy()
        K = self.K
        for c in self.colour_order:
            m = self.maps[c]
            r = rhs[m] - K[m, :] @ x
            x[m] += damp * np.linalg.solve(self.local[c], r)
        return x

    def vanka(self, rhs, x, damp):
        """One CELL_JACOBI sweep: residual once, local solves (LAPACK LU with
        partial pivoting), damped updates averaged over the cells of each DOF."""
        r = rhs - self.K @ x
        upd = np.zeros_like(x)
        for m, L in zip(self.maps, self.local):
            upd[m] += damp * np.linalg.solve(L, r[m])
        nz = self.count > 0
        x = x.copy()
        x[nz] += upd[nz] / self.count[nz]
        return x

    def defect(self, rhs, x):
        return rhs - self.ns.apply(self.A, self.B, x)


def nse_mg_apply(levels, v, damp_corr=1.0, vanka_damp=1.0, n_pre=1, n_post=1, coarse="direct_solve",
                 coarse_maxit=10, coarse_res=0.1, cycle="V", smoother="cell_vanka_jacobi"):
    """Multigrid::cycle (src/Multigrid/Multigrid.C:157-420) for NSE levels
    (coarsest first): V/W cycle of CycleControl, velocity (Q2) and pressure (Q1)
    transfers of GridTransfer block by block, restricted defect with the
    coarse Dirichlet velocity entries reset."""
    nl = len(levels)
    cc = _cycle_control(cycle, nl)
    rhs = [None] * nl
    sol = [None] * nl
    F = levels[-1]
    nu, na = F.ns.n_u, F.ns.n_u_active
    rhs[-1] = v.copy()
    x = np.zeros_like(v)
    for k in range(3):
        x[k * nu + na:(k + 1) * nu] = v[k * nu + na:(k + 1) * nu]
    sol[-1] = x

    def smooth(l):
        L = levels[l]
        kind = coarse if l == 0 else smoother
        if kind == "direct_solve":
            sol[0] = np.linalg.solve(L.K.toarray(), rhs[0])
        elif kind == "cell_vanka":
            sol[l] = L.vanka_mult(rhs[l], sol[l], vanka_damp)
        else:
            sol[l] = L.vanka(rhs[l], sol[l], vanka_damp)

    def up(l):   # coarse l -> fine l + 1
        C, Fn = levels[l], levels[l + 1]
        cu, fu = C.ns.n_u, Fn.ns.n_u
        for k in range(3):
            pf = prolongate(C.ns.vs, Fn.ns.vs, sol[l][k * cu:(k + 1) * cu])
            sol[l + 1][k * fu:k * fu + Fn.ns.n_u_active] += damp_corr * pf[:Fn.ns.n_u_active]
        pp = pressure_prolongate(C.ns, Fn.ns, sol[l][3 * cu:])
        sol[l + 1][3 * fu:] += damp_corr * pp

    def down(l):   # fine l -> coarse l - 1
        Fn, C = levels[l], levels[l - 1]
        d = Fn.defect(rhs[l], sol[l])
        cu, fu = C.ns.n_u, Fn.ns.n_u
        r = np.zeros(C.ns.n)
        for k in range(3):
            rc = restrict(C.ns.vs, Fn.ns.vs, d[k * fu:(k + 1) * fu])
            rc[C.ns.n_u_active:] = 0.0
            r[k * cu:(k + 1) * cu] = rc
        r[3 * cu:] = pressure_restrict(C.ns, Fn.ns, d[3 * fu:])
        rhs[l - 1] = r
        sol[l - 1] = np.zeros(C.ns.n)

    level = nl - 1
    for step in range(1, len(cc)):
        if level == 0:
            if coarse == "direct_solve":
                smooth(0)
            else:
                res = 1e10
                i = 0
                while i < coarse_maxit and res > coarse_res:
                    smooth(0)
                    res = np.linalg.norm(levels[0].defect(rhs[0], sol[0]))
                    i += 1
            if nl > 1:
                up(0)
            level = 1
            continue
        if cc[step - 1] == "up":
            for _ in range(n_post):
                smooth(level)
        if cc[step] == "down":
            for _ in range(n_pre):
                smooth(level)
            down(level)
            level -= 1
        elif cc[step] == "up":
            up(level)
            level += 1
    return sol[-1]


def _cycle_control(cycle, nl):
    """CycleControl (src/Multigrid/CycleControl.C:17-111): step directions,
    entry 0 a dummy, the last one 'end'."""
    rec = [1 if cycle == "V" else 2] * nl
    rec[nl - 1] = 1
    cc = ["dummy"]

    def fill(level):
        if level == 0:
            cc.append("up")
            return
        for _ in range(rec[level]):
            cc.append("down")
            fill(level - 1)
        if cycle == "F":
            rec[level] = 1
        cc.append("end" if level == nl - 1 else "up")
    fill(nl - 1)
    return cc


def fgmres(apply, prec, b, x0, restart=30, max_it=100, tol=1e-8, reduction=0.0, min_it=0):
    """Iteration_gmres::flexible_gmres (src/Solver/Iteration_gmres.C:374-532) with
    IterativeMethod::converged: absolute residual |s_{i+1}|, MGS, Givens;
    convergence counts only after more than min_it iterations
    (include/Solver/IterativeMethod.h:97-98)."""
    x = x0.copy()
    r = b - apply(x)
    beta = np.linalg.norm(r)
    r0 = beta
    if (beta <= tol or beta <= reduction * r0) and min_it == 0:
        return x, 0
    j = 1
    while j <= max_it:
        V = [r / beta]
        Z = []
        H = np.zeros((restart + 1, restart))
        cs, sn, s = np.zeros(restart), np.zeros(restart), np.zeros(restart + 1)
        s[0] = beta
        for i in range(restart):
            Z.append(prec(V[i]))
            w = apply(Z[i])
            for k in range(i + 1):
                H[k, i] = np.dot(w, V[k])
                w = w - H[k, i] * V[k]
            H[i + 1, i] = np.linalg.norm(w)
            V.append(w / H[i + 1, i])
            for k in range(i):
                t = cs[k] * H[k, i] + sn[k] * H[k + 1, i]
                H[k + 1, i] = -sn[k] * H[k, i] + cs[k] * H[k + 1, i]
                H[k, i] = t
            dx, dy = H[i, i], H[i + 1, i]
            if dy == 0:
                cs[i], sn[i] = 1.0, 0.0
            elif abs(dy) > abs(dx):
                t = dx / dy
                sn[i] = 1 / np.sqrt(1 + t * t)
                cs[i] = t * sn[i]
            else:
                t = dy / dx
                cs[i] = 1 / np.sqrt(1 + t * t)
                sn[i] = t * cs[i]
            H[i, i] = cs[i] * H[i, i] + sn[i] * H[i + 1, i]
            H[i + 1, i] = 0.0
            s[i + 1] = -sn[i] * s[i]
            s[i] = cs[i] * s[i]
            res = abs(s[i + 1])
            if ((res <= tol or res <= reduction * r0) and (j > min_it or min_it == 0)) or j >= max_it:
                y = np.linalg.solve(np.triu(H[:i + 1, :i + 1]), s[:i + 1])
                for k in range(i + 1):
                    x = x + y[k] * Z[k]
                return x, j
            j += 1
        y = np.linalg.solve(np.triu(H[:restart, :restart]), s[:restart])
        for k in range(restart):
            x = x + y[k] * Z[k]
        r = b - apply(x)
        beta = np.linalg.norm(r)
    return x, max_it


# ------------------------------------------------ nonlinear (Picard) iteration
def _q2_values(xi):
    """Q2 Lagrange basis on [-1, 1]^3 at points xi [m, 3] -> [m, 27] (node
    order of ref_nodes: x fastest)."""
    def p2(t):
        return np.stack([0.5 * t * (t - 1.0), 1.0 - t * t, 0.5 * t * (t + 1.0)], -1)
    bx, by, bz = p2(xi[:, 0]), p2(xi[:, 1]), p2(xi[:, 2])
    return np.stack([bx[:, k % 3] * by[:, (k // 3) % 3] * bz[:, k // 9] for k in range(27)], -1)


def _gauss3_cube():
    g = np.sqrt(0.6)
    p1, w1 = np.array([-g, 0.0, g]), np.array([5.0, 8.0, 5.0]) / 9.0
    X, Y, Z = np.meshgrid(p1, p1, p1, indexing="ij")
    W = np.einsum("i,j,k->ijk", w1, w1, w1)
    pts = np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], -1)
    return pts, W.ravel(order="F")


def restrict_function(cs, fs, uf):
    """GridTransfer::RestrictFunction (src/Multigrid/GridTransfer.C:1218-1501)
    for Q2 on axis-parallel hexahedra: fine cells are visited in order; the
    first child met of a parent triggers the parent's children (in order),
    Val2 = sum_j G^-1 R_j u_j with G the coarse reference mass matrix and R_j
    the coarse/fine mixed mass matrix of child j (both from
    TFEDatabase3D::GetRestrictionMatrix3D, src/FE/FEDatabase3D.C:1771-1990),
    added to the parent's DOFs; finally every coarse value is divided by the
    number of coarse cells containing it."""
    pts, w = _gauss3_cube()
    phc = _q2_values(pts)
    G = phc.T @ (w[:, None] * phc)
    uc = np.zeros(cs.n_dof)
    aux = np.zeros(cs.n_dof)
    children = {}
    for fc in range(fs.n_cells):
        children.setdefault(int(fs.parent[fc]), []).append(fc)
    handled = set()
    def affine(verts):
        # reference map of the cell (vertex-weighted, affine for parallelepipeds)
        x0 = node_phys(HEXA, verts, np.zeros(3))
        J = np.stack([node_phys(HEXA, verts, e) - x0 for e in np.eye(3)], -1)
        return x0, J

    for fc in range(fs.n_cells):
        pc = int(fs.parent[fc])
        if pc in handled:
            continue
        handled.add(pc)
        p0, pJ = affine(cs.verts[pc])
        val2 = np.zeros(27)
        for ch in children[pc]:
            c0, cJ = affine(fs.verts[ch])
            x = c0 + pts @ cJ.T                                    # child reference -> physical
            xp = np.linalg.solve(pJ, (x - p0).T).T                 # physical -> parent reference
            R = _q2_values(xp).T @ ((w / 8.0)[:, None] * phc)      # R[k, l] = (phi_k^H, phi_l^h)
            RQ = np.linalg.solve(G, R)
            val2 += RQ @ uf[fs.gnum[ch]]
        dofs = cs.gnum[pc]
        uc[dofs] += val2
        aux[dofs] += 1.0
    return uc / aux


def q1_mean(ps, p):
    """Mean value of a Q1 function on the union of its (axis-parallel
    hexahedral) cells: a trilinear function's integral over a brick is the
    brick volume times the mean of its vertex values."""
    vol = np.prod(ps.verts.max(1) - ps.verts.min(1), axis=1)
    return float(np.sum(vol * p[ps.gnum].mean(1)) / np.sum(vol))


def nse_picard(n_refine, levels, nu, mg_kw, lin_tol=1e-11, lin_max_it=200, restart=30, eps=1e-10, max_it=10,
               min_it=1, slowfactor=1e10, pressure="q1", example="u2_p1", lin_min_it=0):
    """NavierStokes<3> nonlinear loop as test/nse3d_solvers_test.c++:107-126
    runs it: assemble_linear_terms (convection with the initial, zero,
    velocity; Dirichlet values copied into the solution), stop_it(0), then
    repeatedly solve (FGMRES + saddle multigrid from the current solution,
    pressure projected into L2_0), assemble_nonlinear_term (velocity restricted
    to the coarser levels, GridTransfer::RestrictFunctionRepeatedly) and
    stop_it(k) (src/System/NavierStokes.C:750-818: full residual
    sqrt(|r_u|^2 + |r_p|^2) of rhs - K x with the pressure part of the defect
    projected into L2_0).  Returns the solution, the residual history and the
    FGMRES iteration counts.  The pressure is projected only for an enclosed
    flow (BlockFEMatrix::pressure_projection_enabled)."""
    sys_ = [NseOracle(n_refine, l, pressure, example) for l in levels]
    sols = [np.zeros(s.n) for s in sys_]
    F = sys_[-1]
    nu_, na = F.n_u, F.n_u_active

    def assemble():
        return [s.assemble(nu, x[:3 * s.n_u]) for s, x in zip(sys_, sols)]

    blocks = assemble()
    rhs = blocks[-1][2]
    x = sols[-1]
    for k in range(3):
        x[k * nu_ + na:(k + 1) * nu_] = rhs[k * nu_ + na:(k + 1) * nu_]

    def residual(A, B):
        d = rhs - F.apply(A, B, x)
        if F.enclosed:
            d[3 * nu_:] -= F.pressure_mean(d[3 * nu_:])
        return np.sqrt(np.dot(d[:3 * nu_], d[:3 * nu_]) + np.dot(d[3 * nu_:], d[3 * nu_:]))

    hist = [residual(blocks[-1][0], blocks[-1][1])]
    lin_its = []
    k = 0
    while True:
        k += 1
        lv = [NseLevelOracle(s, b[0], b[1]) for s, b in zip(sys_, blocks)]
        A, B = blocks[-1][0], blocks[-1][1]
        x, its = fgmres(lambda z: F.apply(A, B, z), lambda v: nse_mg_apply(lv, v, **mg_kw), rhs, x,
                        restart=restart, max_it=lin_max_it, tol=lin_tol, min_it=lin_min_it)
        if F.enclosed:
            x[3 * nu_:] -= F.pressure_mean(x[3 * nu_:])
        lin_its.append(its)
        sols[-1] = x
        for l in range(len(sys_) - 2, -1, -1):
            cs, fsv = sys_[l], sys_[l + 1]
            for c in range(3):
                sols[l][c * cs.n_u:(c + 1) * cs.n_u] = restrict_function(
                    cs.vs, fsv.vs, sols[l + 1][c * fsv.n_u:(c + 1) * fsv.n_u])
        blocks = assemble()
        hist.append(residual(blocks[-1][0], blocks[-1][1]))
        old = hist[max(0, len(hist) - 10)]
        if k >= min_it and (hist[-1] <= eps or k == max_it or hist[-1] >= slowfactor * old):
            break
    return x, hist, lin_its
