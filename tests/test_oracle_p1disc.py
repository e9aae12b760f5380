"""Discontinuous P1 pressure (Q2/P1disc, the pair of the reference's cell-Vanka
test) in the oracle, pinned by what the mathematics and the reference's test
fix:

* the prolongation of a coarse P1disc function is that function on the fine
  cells (P1 restricted to a child is P1; every fine DOF lies in one cell);
* the divergence blocks on velocities in the Q2 space: (B1 u)_i =
  -int psi_i du/dx per cell, so for u = x the constant row is -|K| and the xi
  rows vanish, for u = x^2 the row of the reference axis along x is
  -|K| h / 3 (h the cell's edge along x), signed by the cell's orientation;
* the velocity space of test_u_0_p_0 (Neumann on the interior of x = 1):
  the open face's DOFs are active, its rim stays Dirichlet;
* the reference's own test (test/nse3d_solvers_test.c++:334-341 with the
  cell_vanka_jacobi settings of :171-191): example -1, Q2/P1disc, two levels,
  FGMRES + V(1,1) cell Vanka, Picard to 1e-12: errors below its tolerance 1e-8.
All of it runs on the CPU."""
import numpy as np

from oracle import oracle as O


def _cell_geometry(space):
    lo = space.verts.min(1)
    hi = space.verts.max(1)
    return lo, hi


def test_p1disc_prolongation_reproduces_coarse_function():
    c = O.NseOracle(2, 1, "p1disc", "u0_p0")
    f = O.NseOracle(2, 2, "p1disc", "u0_p0")
    P = O.p1disc_prolongation(c.ps, f.ps)
    rng = np.random.default_rng(11)
    pc = rng.standard_normal(c.n_p)
    pf = P @ pc

    def value(space, coeffs, cell, x):
        x0, J = O._affine(space.verts[cell])
        xi = np.linalg.solve(J, x - x0)
        return coeffs[space.gnum[cell]] @ np.array([1.0, xi[0], xi[1], xi[2]])

    for fc in range(f.ps.n_cells):
        lo, hi = _cell_geometry(f.ps)
        for _ in range(3):
            x = lo[fc] + (hi[fc] - lo[fc]) * rng.random(3)
            a = value(c.ps, pc, int(f.ps.parent[fc]), x)
            b = value(f.ps, pf, fc, x)
            assert abs(a - b) <= 1e-13 * max(1.0, abs(a)), (fc, a, b)
    # every fine DOF has one row of at most 4 coarse entries from its parent
    assert np.all(np.diff(P.indptr) <= 4)


def test_p1disc_divergence_blocks_closed_forms():
    o = O.NseOracle(2, 2, "p1disc", "u2_p1")
    A, B, rhs = o.assemble(1.0)
    import scipy.sparse as sp
    Bm = [sp.csr_matrix((b, o.bcol, o.brp), shape=(o.n_p, o.n_u)) for b in B]
    xyz = O.nodal_coords(o.vs)
    lo, hi = _cell_geometry(o.ps)
    vol = np.prod(hi - lo, axis=1)
    for k in range(3):
        # u_k = x_k: B_k u = -int psi_i
        g = Bm[k] @ xyz[:, k]
        want = np.zeros(o.n_p)
        want[o.ps.gnum[:, 0]] = -vol
        assert np.max(np.abs(g - want)) <= 1e-13
        # u_k = x_k^2 = (x0_k + sum_m J_km xi_m)^2: the constant row is
        # -2 x0_k |K| and the xi_m row -2 J_km |K| / 3 (mean of xi_m^2 = 1/3),
        # i.e. -|K| h / 3 on the reference axis along x_k, signed by the
        # cell's orientation (RefHexaRegDesc's children are not all oriented
        # alike)
        g = Bm[k] @ (xyz[:, k] ** 2)
        want = np.zeros(o.n_p)
        n_axis = 0
        for c in range(o.ps.n_cells):
            x0, J = O._affine(o.ps.verts[c])
            want[o.ps.gnum[c, 0]] = -vol[c] * 2 * x0[k]
            for m in range(3):
                want[o.ps.gnum[c, 1 + m]] = -2 * J[k, m] * vol[c] / 3
            n_axis += int(np.sum(np.abs(J[k]) > 1e-12))
        assert n_axis == o.ps.n_cells   # one reference axis along x_k per brick
        assert np.max(np.abs(g - want)) <= 1e-13, k
        # other components do not see x_k^2
        for m in range(3):
            if m != k:
                assert np.max(np.abs(Bm[m] @ (xyz[:, k] ** 2))) <= 1e-13


def test_u0p0_velocity_space_opens_the_face_x1():
    o = O.NseOracle(2, 2, "q1", "u0_p0")
    xyz = O.nodal_coords(o.vs)
    on_face = np.abs(xyz[:, 0] - 1) < 1e-12
    rim = on_face & ((np.abs(xyz[:, 1]) < 1e-12) | (np.abs(xyz[:, 1] - 1) < 1e-12) |
                     (np.abs(xyz[:, 2]) < 1e-12) | (np.abs(xyz[:, 2] - 1) < 1e-12))
    active = np.arange(o.n_u) < o.n_u_active
    assert np.all(active[on_face & ~rim]) and not np.any(active[rim])
    n = 9   # 4^3 cells: 9 x 9 Q2 nodes on the face, 7 x 7 inside
    assert int(np.sum(on_face & ~rim)) == (n - 2) ** 2
    assert not o.enclosed
    assert O.NseOracle(2, 2, "q1", "u2_p1").enclosed


def test_reference_cell_vanka_test_on_the_oracle():
    """test/nse3d_solvers_test.c++ cell_vanka_jacobi: example -1 on the twice
    refined unit cube (4^3 hexahedra), Q2/P1disc, 2 multigrid levels, FGMRES
    (residual 1e-8, at most 10, at least 3 iterations), V(1,1) with
    cell_vanka_jacobi smoother and coarse smoother (damping 0.8, coarse
    residual 0.1 within 10 sweeps), correction damping 0.8, Picard to 1e-12
    within 15 steps; the expected errors are 0 with tolerance 1e-8."""
    mg = dict(damp_corr=0.8, vanka_damp=0.8, n_pre=1, n_post=1, coarse="cell_vanka_jacobi", coarse_maxit=10,
              coarse_res=0.1, cycle="V", smoother="cell_vanka_jacobi")
    x, hist, its = O.nse_picard(2, [1, 2], 1.0, mg, lin_tol=1e-8, lin_max_it=10, restart=20, eps=1e-12, max_it=15,
                                pressure="p1disc", example="u0_p0", lin_min_it=3)
    F = O.NseOracle(2, 2, "p1disc", "u0_p0")
    exact = np.zeros(F.n)
    exact[:F.n_u] = 1.0
    assert hist[-1] <= 1e-12 and len(hist) <= 16, hist
    assert all(i >= 4 for i in its), its   # more than min_n_iterations = 3
    assert np.max(np.abs(x - exact)) < 1e-8
