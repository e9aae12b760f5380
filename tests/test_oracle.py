"""Pins the CPU oracle (oracle/) against the reference's own golden values and
against closed forms, before it is used as the checker of the HIP path."""
import numpy as np
import pytest

from conftest import iterative_7x7
from oracle import oracle as O

GEO = {"UnitSquare": 0, "TwoTriangles": 1, "Default_UnitCube_Hexa": 2}
EX = {"const": 0, "sine_laplace_2d": 1, "p0_3d": 2, "p1_3d": 3, "p2_3d": 4, "exp_sin_3d": 5, "sine_2d": 6}


def solve_direct_like(space, val, rhs, tol=1e-14):
    """FGMRES without preconditioner run to round-off (stands in for the
    reference's direct solver in the golden tests)."""
    x0 = np.zeros(space.n_dof)
    x0[space.active:] = rhs[space.active:]
    sp = O.solver_params(iterative_solver_type=0, max_n_iterations=3000, gmres_restart=400,
                         residual_tolerance=tol)
    x, it, hist, st = O.solve(sp, [space], [val], rhs, x0)
    assert st == 0
    return x


@pytest.mark.parametrize("geo,order", [("UnitSquare", 1), ("UnitSquare", 2), ("TwoTriangles", 1)])
def test_cd2d_golden_errors(golden, geo, order):
    ref = golden["cd2d_test"][geo][str(order)]
    tol = golden["cd2d_test"]["tolerance"]
    sp, info = O.build_space(GEO[geo], 2, 2, order, cond2d=lambda c, t: 1 if c == 1 else 0)
    p = O.cd_params(example=EX["sine_laplace_2d"], sine_laplace_bc=True)
    val, rhs = O.assemble(sp, p)
    x = solve_direct_like(sp, val, rhs)
    err = O.errors(sp, p, x)
    assert np.all(np.abs(err - np.array(ref)) < tol), (err, ref)


def test_cd2d_golden_detects_planted_error(golden):
    """A dropped Neumann term (planted in the oracle inputs) breaks the golden check."""
    ref = golden["cd2d_test"]["UnitSquare"]["1"]
    sp, info = O.build_space(0, 2, 2, 1, cond2d=lambda c, t: 1 if c == 1 else 0)
    p = O.cd_params(example=EX["sine_laplace_2d"], sine_laplace_bc=False)   # Neumann integral dropped
    val, rhs = O.assemble(sp, p)
    x = solve_direct_like(sp, val, rhs)
    err = O.errors(sp, p, x)
    assert abs(err[0] - ref[0]) > 1e-6


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_cd3d_zero_error_examples(golden, idx):
    name, order, eps, b, c = golden["cd3d_solvers_test"]["examples"][idx]
    tol = golden["cd3d_solvers_test"]["tolerance_direct"]
    sp, info = O.build_space(2, 2, 2, order)
    p = O.cd_params(example=EX[name], eps=eps, b=b, c=c)
    val, rhs = O.assemble(sp, p)
    x = solve_direct_like(sp, val, rhs, tol=1e-15)
    err = O.errors(sp, p, x)
    # reference tolerance 1e-13 for the direct solver; our Krylov stand-in
    # reaches the same level of round-off
    assert err[0] < 10 * tol and err[1] < 1e3 * tol, err


def test_matrix_test_spmv_golden(golden):
    g = golden["matrix_test"]

    def dense(m):
        a = np.zeros((m["n_rows"], m["n_cols"]))
        for r in range(m["n_rows"]):
            for k in range(m["row_ptr"][r], m["row_ptr"][r + 1]):
                a[r, m["col"][k]] = m["val"][k]
        return a

    CB = dense(g["C"]) @ dense(g["B"])
    rp = [0]
    col, val = [], []
    for r in range(4):
        nz = np.nonzero(CB[r])[0]
        col += list(nz)
        val += list(CB[r, nz])
        rp.append(len(col))
    y = np.zeros(4)
    O.apply_scaled_add(4, 4, rp, col, val, np.array(g["x"]), y, -0.5)
    ref = np.array(g["y_minus_half_CB_x"])
    assert np.all(np.abs((y - ref) / ref) < 1e-14), (y, ref)


@pytest.mark.parametrize("method,prec", [(3, 0), (2, 1), (1, 1), (0, 1), (0, 0)])
def test_iterative_solver_7x7(golden, method, prec):
    g = golden["iterative_solver_test"]
    M, rp, col, val = iterative_7x7(g)
    exact = np.arange(7, dtype=float)
    b = M @ exact
    sp = O.csr_general(7, 7, rp, col)
    params = O.solver_params(iterative_solver_type=method, preconditioner=prec,
                             residual_tolerance=g["residual_tolerance"],
                             max_n_iterations=g["max_n_iterations"])
    x, it, hist, st = O.solve(params, [sp], [val], b, np.zeros(7))
    assert st == 0
    assert np.linalg.norm(x - exact) < 100 * g["residual_tolerance"]


def one_cell_space(order):
    """single hexahedron [0,1]^3, every row active (no boundary joints)"""
    shape, levels = O.mesh(2, 0)
    nd = O.ndof_cell(shape, order)
    g = np.arange(nd, dtype=np.int32).reshape(1, nd)
    rp, col = O.structure(g, nd, nd)
    return O.OracleSpace(shape, order, levels[0]["verts"], g, nd, nd, rp, col,
                         np.full((1, 6), -1, np.int32))


def to_dense(sp, val):
    A = np.zeros((sp.n_dof, sp.n_dof))
    for r in range(sp.n_dof):
        A[r, sp.col[sp.rp[r]:sp.rp[r + 1]]] = val[sp.rp[r]:sp.rp[r + 1]]
    return A


def test_q2_brick_laplacian_is_kronecker_sum():
    """On [0,1]^3 the Q2 Laplace matrix is Kx(M)M + M(x)K(x)M + M(x)M(x)K with the
    1D P2 stiffness/mass matrices (dof index = ix + 3 iy + 9 iz)."""
    sp = one_cell_space(2)
    val, rhs = O.assemble(sp, O.cd_params(example=0, eps=1.0, c=0.0))
    A = to_dense(sp, val)
    K = np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / 3.0
    M = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) / 30.0
    ref = np.kron(np.kron(M, M), K) + np.kron(np.kron(M, K), M) + np.kron(np.kron(K, M), M)
    assert np.max(np.abs(A - ref)) < 1e-13


def test_mass_matrix_sums_to_volume_and_convection_rows_vanish():
    sp = one_cell_space(2)
    val, _ = O.assemble(sp, O.cd_params(example=0, eps=0.0, c=1.0))
    assert abs(val.sum() - 1.0) < 1e-13
    val, _ = O.assemble(sp, O.cd_params(example=0, eps=0.0, b=(0.3, -1.1, 2.0)))
    A = to_dense(sp, val)
    assert np.max(np.abs(A.sum(axis=1))) < 1e-13   # b.grad(1) = 0


def test_prolongation_reproduces_polynomials_and_restriction_is_adjoint():
    coarse, _ = O.build_space(2, 2, 1, 2)
    fine, _ = O.build_space(2, 2, 2, 2)
    sh, levels = O.mesh(2, 2)

    def nodal(space, level, f):
        u = np.zeros(space.n_dof)
        xi = O.ref_nodes(sh, 2)
        for c in range(space.n_cells):
            for k in range(space.ndof_cell):
                x = O.node_phys(sh, levels[level]["verts"][c], xi[k])
                u[space.gnum[c, k]] = f(x)
        return u

    f = lambda x: 1 + x[0] - 2 * x[1] * x[2] + 3 * x[0] ** 2   # noqa: E731 (in Q2)
    uc = nodal(coarse, 1, f)
    uf = O.prolongate(coarse, fine, uc)
    assert np.max(np.abs(uf - nodal(fine, 2, f))) < 1e-13
    rng = np.random.default_rng(1)
    u = rng.standard_normal(coarse.n_dof)
    d = rng.standard_normal(fine.n_dof)
    lhs = O.prolongate(coarse, fine, u) @ d
    rhs = u @ O.restrict(coarse, fine, d)
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_multigrid_fgmres_solves_quadratic_example():
    """cd3d_solvers_test 'multigrid' setup: FGMRES + MG(W, 2 levels) on example -3."""
    sc, _ = O.build_space(2, 2, 1, 2)
    sf, _ = O.build_space(2, 2, 2, 2)
    p = O.cd_params(example=EX["p2_3d"], eps=1.0, b=(4, 3, 2), c=1.0)
    vc, rc = O.assemble(sc, p)
    vf, rf = O.assemble(sf, p)
    x0 = np.zeros(sf.n_dof)
    x0[sf.active:] = rf[sf.active:]
    params = O.solver_params(iterative_solver_type=0, preconditioner=2, max_n_iterations=100,
                             residual_tolerance=1e-15, mg_cycle=1, mg_smoother_coarse=1,
                             mg_correction_damp=0.8, mg_n_pre_smooth=3, mg_n_post_smooth=3)
    x, it, hist, st = O.solve(params, [sc, sf], [vc, vf], rf, x0)
    err = O.errors(sf, p, x)
    assert err[0] < 1e-12 and err[1] < 1e-12
    assert it < 60


@pytest.mark.parametrize("disc", ["galerkin", "supg"])
def test_tcd2d_golden_errors(golden, disc):
    """test/tcd2d_test.c++: Crank-Nicolson, exp example, Q1 on 32x32 quads;
    L2/H1 errors after the listed time steps."""
    g = golden["tcd2d_test"]
    sp, info = O.build_space(0, 5, 5, 1)
    errs, _ = O.tcd_time_integration(sp, 7, 1 if disc == "supg" else 0, O.CRANK_NICOLSON,
                                     g["time_step_length"], g["n_steps"], eps=1.0, b=(1.0, -1.0), c=1.0)
    for k, m in enumerate(g["steps"]):
        assert abs(errs[m][0] - g[disc]["L2"][k]) < g["tolerance"], (m, errs[m][0], g[disc]["L2"][k])
        assert abs(errs[m][1] - g[disc]["H1"][k]) < g["tolerance"], (m, errs[m][1], g[disc]["H1"][k])


def test_tcd2d_golden_detects_planted_error(golden):
    """Backward Euler in place of Crank-Nicolson (planted in the oracle run)
    misses the golden errors."""
    g = golden["tcd2d_test"]
    sp, info = O.build_space(0, 5, 5, 1)
    errs, _ = O.tcd_time_integration(sp, 7, 0, O.BACKWARD_EULER, g["time_step_length"], 3,
                                     eps=1.0, b=(1.0, -1.0), c=1.0)
    assert abs(errs[3][0] - g["galerkin"]["L2"][3]) > 1e-6


@pytest.mark.parametrize("scheme", [0, 1, 2])
@pytest.mark.parametrize("geo,order", [("UnitSquare", 1), ("TwoTriangles", 1), ("UnitSquare", 2)])
def test_tcd_linear_space_time_is_exact(scheme, geo, order):
    """TCD_2D/linear_space_time.h (u = 1 + 2x + 3ty): the Galerkin space
    contains u(t) and every scheme integrates u_t = 3y exactly, so the
    discrete solution equals the interpolant at every step."""
    sp, info = O.build_space(GEO[geo], 3, 3, order)
    errs, _ = O.tcd_time_integration(sp, 8, 0, scheme, 0.1, 5, eps=1.0, b=(1.0, 2.0), c=1.0)
    assert np.max(errs) < 1e-10, errs


@pytest.mark.parametrize("nref,nu", [(1, 1.0), (2, 1e-2)])
def test_nse_interpolant_solves_oseen_with_its_own_convection(nref, nu):
    """test_u_2_p_1 (NSE3D example -3): u quadratic, p linear, div u = 0, so the
    Q2/Q1 interpolant satisfies the discrete Oseen system with w = u exactly
    (Gauss4 integrates every term exactly): K x_I = rhs to round-off.  The same
    check without the convection term fails, and so does a flipped B sign."""
    ns = O.NseOracle(nref, nref)
    xI = O.nse_interpolate(ns)
    w = xI[:3 * ns.n_u]
    A, B, rhs = ns.assemble(nu, w)
    assert np.max(np.abs(ns.apply(A, B, xI) - rhs)) < 1e-13
    A0, B0, rhs0 = ns.assemble(nu, None)
    assert np.max(np.abs(ns.apply(A0, B0, xI) - rhs0)) > 1e-4
    Bf = [-b for b in B]
    assert np.max(np.abs(ns.apply(A, Bf, xI) - rhs)) > 1e-4


def test_nse_pattern_closed_forms():
    """B rows (pressure vertices) couple to the Q2 DOFs of the cells around the
    vertex: an inner vertex of an n^3 mesh touches 8 cells = 5^3 Q2 DOFs."""
    ns = O.NseOracle(2, 2)
    n = 4
    assert ns.n_u == (2 * n + 1) ** 3 and ns.n_u_active == (2 * n - 1) ** 3 and ns.n_p == (n + 1) ** 3
    lens = np.diff(ns.brp)
    assert lens.max() == 125 and np.sum(lens == 125) == (n - 1) ** 3


def test_function_restriction_closed_forms():
    """GridTransfer::RestrictFunction restated: on nested Q2 spaces the local
    L2 projection returns a prolongated coarse function exactly (RF P = I) and
    reproduces quadratic interpolants; a wrong local DOF order (planted: two
    local DOFs of one fine cell exchanged) breaks RF P = I."""
    cs, fs = O.NseOracle(2, 1).vs, O.NseOracle(2, 2).vs
    uc = np.random.default_rng(1).standard_normal(cs.n_dof)
    uf = O.prolongate(cs, fs, uc)
    assert np.max(np.abs(O.restrict_function(cs, fs, uf) - uc)) < 1e-13
    xf, xc = O.nodal_coords(fs), O.nodal_coords(cs)
    f = lambda x: x[:, 0] ** 2 - 2 * x[:, 1] * x[:, 2] + x[:, 2]
    assert np.max(np.abs(O.restrict_function(cs, fs, f(xf)) - f(xc))) < 1e-13
    # planted: swap two velocity DOFs of one fine cell's local order
    bad = fs.gnum.copy()
    fs2 = O.OracleSpace(fs.shape, fs.order, fs.verts, bad, fs.n_dof, fs.active, fs.rp, fs.col, fs.bd, fs.parent)
    fs2.gnum[0, [0, 1]] = fs2.gnum[0, [1, 0]]
    assert np.max(np.abs(O.restrict_function(cs, fs2, uf) - uc)) > 1e-3


def test_nse_picard_oracle_converges_to_interpolant():
    """test_u_2_p_1 is quadratic/linear, so its interpolant is the discrete
    Navier-Stokes solution (convection included): the Picard iteration
    converges to it, and the nonlinear residual decreases monotonically."""
    kw = dict(damp_corr=0.8, vanka_damp=0.5, n_pre=2, n_post=2)
    x, hist, its = O.nse_picard(2, [1, 2], 0.1, kw, eps=1e-10, max_it=20)
    ns = O.NseOracle(2, 2)
    xi = O.nse_interpolate(ns)
    xi[3 * ns.n_u:] -= O.q1_mean(ns.ps, xi[3 * ns.n_u:])
    assert hist[-1] <= 1e-10
    assert all(b < a for a, b in zip(hist, hist[1:]))
    assert np.max(np.abs(x - xi)) < 1e-8


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_gls_keeps_the_zero_error_examples(golden, idx):
    """GLS is consistent (its terms vanish for the exact solution), so the
    cd3d_solvers_test examples with polynomial solutions in the discrete
    space stay exact; a sign error in the GLS right-hand side (planted on the
    quadratic example) is visible."""
    name, order, eps, b, c = golden["cd3d_solvers_test"]["examples"][idx]
    tol = golden["cd3d_solvers_test"]["tolerance_direct"]
    sp, info = O.build_space(2, 2, 2, order)
    p = O.cd_params(example=EX[name], disc=2, eps=eps, b=b, c=c)
    val, rhs = O.assemble(sp, p)
    x = solve_direct_like(sp, val, rhs, tol=1e-15)
    err = O.errors(sp, p, x)
    assert err[0] < 10 * tol and err[1] < 1e3 * tol, err
    if idx == 2:
        pg = O.cd_params(example=EX[name], disc=0, eps=eps, b=b, c=c)
        vg, rg = O.assemble(sp, pg)
        bad = rhs - 2.0 * (rhs - rg)          # GLS rhs term with the wrong sign
        bad[sp.active:] = rhs[sp.active:]
        xb = solve_direct_like(sp, val, bad, tol=1e-15)
        assert O.errors(sp, p, xb)[0] > 1e-6


def test_gls_term_is_symmetric_and_vanishes_for_q1_diffusion():
    """The GLS term delta (L phi_j, L phi_i) is symmetric (unlike the SUPG term
    delta (L phi_j, b.grad phi_i)), and on Q1 bricks without reaction
    (lap phi = 0, c = 0, so L phi = b.grad phi) GLS equals SUPG."""
    sp, info = O.build_space(2, 2, 2, 2)
    kw = dict(example=EX["exp_sin_3d"], eps=1e-2, b=(1.0, 2.0, 3.0), c=0.5)
    vg, _ = O.assemble(sp, O.cd_params(disc=0, **kw))
    vl, _ = O.assemble(sp, O.cd_params(disc=2, **kw))
    vs, _ = O.assemble(sp, O.cd_params(disc=1, **kw))
    D = to_dense(sp, vl - vg)[:sp.active, :sp.active]
    S = to_dense(sp, vs - vg)[:sp.active, :sp.active]
    assert np.max(np.abs(D - D.T)) < 1e-12 * np.max(np.abs(D))
    assert np.max(np.abs(S - S.T)) > 1e-3 * np.max(np.abs(S))
    sp1, _ = O.build_space(2, 2, 2, 1)
    k1 = dict(example=EX["exp_sin_3d"], eps=1e-2, b=(1.0, 2.0, 3.0), c=0.0)
    a, fa = O.assemble(sp1, O.cd_params(disc=1, **k1))
    g, fg = O.assemble(sp1, O.cd_params(disc=2, **k1))
    ga, _ = O.assemble(sp1, O.cd_params(disc=0, **k1))
    assert np.max(np.abs(a - ga)) > 1e-6
    assert np.max(np.abs(a - g)) < 1e-14 * np.max(np.abs(a))
    assert np.max(np.abs(fa - fg)) < 1e-14 * np.max(np.abs(fa))


def test_multiplicative_vanka_colours_are_independent():
    """The colour-major order of the multiplicative cell Vanka: within a colour
    (cell lattice index mod 3 per direction) no cell reads a DOF another one
    updates, so reversing the order inside every colour changes nothing (the
    GPU runs a colour's cells in parallel); a parity (mod 2) colouring does not
    have this property."""
    o = O.NseOracle(3, 3)
    xo = O.nse_interpolate(o)
    A, B, ro = o.assemble(1e-2, xo[:3 * o.n_u])
    L = O.NseLevelOracle(o, A, B)
    rng = np.random.default_rng(4)
    rhs, x = rng.standard_normal(o.n), rng.standard_normal(o.n)
    a = L.vanka_mult(rhs, x, 0.7)
    ctr = o.vs.verts.mean(1)
    h = o.vs.verts[0].max(0) - o.vs.verts[0].min(0)
    ijk = np.floor(ctr / h).astype(int)
    col3 = (ijk % 3) @ np.array([1, 3, 9])
    col2 = (ijk % 2) @ np.array([1, 2, 4])
    base = L.colour_order
    L.colour_order = [c for k in range(27) for c in sorted(np.where(col3 == k)[0], reverse=True)]
    b = L.vanka_mult(rhs, x, 0.7)
    assert np.max(np.abs(a - b)) < 1e-14 * np.max(np.abs(a))
    L.colour_order = [c for k in range(8) for c in sorted(np.where(col2 == k)[0])]
    p1 = L.vanka_mult(rhs, x, 0.7)
    L.colour_order = [c for k in range(8) for c in sorted(np.where(col2 == k)[0], reverse=True)]
    p2 = L.vanka_mult(rhs, x, 0.7)
    assert np.max(np.abs(p1 - p2)) > 1e-8 * np.max(np.abs(p1))
    L.colour_order = base
