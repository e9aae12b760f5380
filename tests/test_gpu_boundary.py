"""The C-ABI boundary as a caller meets it: argument checks that run before
any kernel, the error codes the reference turns into exceptions (zero
diagonal, divergence, unhappy breakdown), the BLAS-1 entry points, the
F-cycle, the CD2D Poisson configuration (64 x 64 Q1, CG + Jacobi to 1e-10),
the three-component SELL pass of the saddle-point operator, and the two
routes for caller data: a coefficient table in place of the CoeffFct
callback and caller-given Dirichlet values."""
import math

import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from mymoonmd_amd._lib import PmxError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

B_CD3D = (1 / 14 ** 0.5, 2 / 14 ** 0.5, 3 / 14 ** 0.5)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def csr_of(M):
    rp, col, val = [0], [], []
    for r in range(M.shape[0]):
        nz = np.nonzero(M[r])[0]
        col += list(nz)
        val += list(M[r, nz])
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val)


def test_vector_arguments_are_checked_before_any_kernel(ctx):
    """A host tensor, a float32 tensor, a strided view or a short tensor is
    refused with ValueError; nothing is launched (the matrix stays usable)."""
    rp, col, val = csr_of(np.diag([2.0, 3.0, 4.0, 5.0]))
    M = pm.FEMatrix(ctx, csr=(rp, col, val, 4))
    y = torch.zeros(4, dtype=torch.float64, device="cuda")
    bad = [torch.ones(4, dtype=torch.float64),                          # host memory
           torch.ones(4, dtype=torch.float32, device="cuda"),           # wrong type
           torch.ones(8, dtype=torch.float64, device="cuda")[::2],      # not contiguous
           torch.ones(3, dtype=torch.float64, device="cuda"),           # too short
           np.ones(4)]                                                  # not a tensor
    for x in bad:
        with pytest.raises(ValueError):
            M.apply(x, y)
        with pytest.raises(ValueError):
            M.apply_scaled_add(y, x if torch.is_tensor(x) else y[:3])
    sol = pm.Solver(ctx, pm.solver_params(iterative_solver_type="cg", preconditioner="jacobi"))
    with pytest.raises(ValueError):
        sol.solve(y, y)                     # not updated: the length is unknown
    sol.update_matrix(M)
    with pytest.raises(ValueError):
        sol.solve(y[:3], y)
    x = torch.ones(4, dtype=torch.float64, device="cuda")
    M.apply(x, y)
    assert y.cpu().numpy().tolist() == [2.0, 3.0, 4.0, 5.0]


def test_zero_diagonal_is_reported(ctx):
    """Iteration_jacobi's check (src/Solver/Iteration_jacobi.C:8-15) as PMX_E_ZERO_DIAG."""
    A = np.array([[2.0, 1.0, 0.0], [1.0, 0.0, 1.0], [0.0, 1.0, 2.0]])
    rp, col, val = csr_of(A)
    M = pm.FEMatrix(ctx, csr=(rp, col, val, 3))
    for method, prec in (("cg", "jacobi"), ("jacobi", "no_preconditioner")):
        sol = pm.Solver(ctx, pm.solver_params(iterative_solver_type=method, preconditioner=prec))
        with pytest.raises(PmxError) as e:
            sol.update_matrix(M)
        assert e.value.code == 3 and "zero entry on the diagonal" in str(e.value)


def test_divergence_is_reported(ctx):
    """IterativeMethod::converged throws when the residual exceeds
    divergence_factor times the initial one (include/Solver/IterativeMethod.h:
    105-131): Richardson with damping 3 on the identity doubles the residual
    every step (x_k - x = (-2)^k (x_0 - x))."""
    n = 16
    rp, col, val = csr_of(np.eye(n))
    M = pm.FEMatrix(ctx, csr=(rp, col, val, n))
    b = dev(np.linspace(1.0, 2.0, n))
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    sol = pm.Solver(ctx, pm.solver_params(iterative_solver_type="richardson", preconditioner="no_preconditioner",
                                          damping=3.0, divergence_factor=100.0, max_n_iterations=50,
                                          residual_tolerance=1e-12))
    sol.update_matrix(M)
    with pytest.raises(PmxError) as e:
        sol.solve(b, x)
    assert e.value.code == 4 and "divergence" in str(e.value)
    # 2^k > 100 first at k = 7: the history holds the initial and 7 residuals
    h = sol.history()
    assert len(h) == 8 and h[-1] == pytest.approx(128.0 * h[0], rel=1e-12)


def test_unhappy_breakdown_is_reported(ctx):
    """Iteration_gmres.C:470-477: H(i+1, i) == 0 before the residual test
    throws.  A = 2 I and b = e_0: v_0 = e_0, A v_0 = 2 e_0, H(0,0) = 2 exactly
    and the orthogonalised vector is exactly zero."""
    n = 8
    rp, col, val = csr_of(2.0 * np.eye(n))
    M = pm.FEMatrix(ctx, csr=(rp, col, val, n))
    b = np.zeros(n)
    b[0] = 1.0
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    sol = pm.Solver(ctx, pm.solver_params(iterative_solver_type="fgmres", preconditioner="no_preconditioner",
                                          residual_tolerance=1e-300, max_n_iterations=10))
    sol.update_matrix(M)
    with pytest.raises(PmxError) as e:
        sol.solve(dev(b), x)
    assert e.value.code == 5 and "Unhappy breakdown" in str(e.value)


def test_dot_and_norm(ctx):
    """BlockVector::dot / norm (Ddot, Dnorm of src/FE/Blas1.C:14-117: sequential
    sums) against exactly rounded values: the device sums in another order, so
    the bound is n eps times the sum of magnitudes."""
    rng = np.random.default_rng(11)
    for n in (1, 3, 1000, 65537, 2_000_003):
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        exact_dot = math.fsum(x * y)
        exact_nrm = math.sqrt(math.fsum(x * x))
        gd = pm.vector_dot(ctx, dev(x), dev(y))
        gn = pm.vector_norm(ctx, dev(x))
        bound = n * np.finfo(float).eps * float(np.sum(np.abs(x * y)))
        assert abs(gd - exact_dot) <= bound, n
        assert abs(gn - exact_nrm) <= n * np.finfo(float).eps * exact_nrm, n
        yy = dev(y)
        pm.vector_add_scaled(ctx, dev(x), yy, -0.25)
        assert np.array_equal(yy.cpu().numpy(), y + (-0.25) * x)
    assert pm.vector_norm(ctx, dev(np.zeros(17))) == 0.0
    assert pm.vector_dot(ctx, dev([1.5, -2.0]), dev([4.0, 0.25])) == 5.5


def three_level_cd3d(ctx, nref, kw):
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    spaces = [pm.TFESpace(d, l, 2, ctx) for l in range(nref - 2, nref + 1)]
    mats, rhs = [], []
    for s in spaces:
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.assemble_cd(ctx, s, pm.cd_params(**kw), A, f)
        mats.append(A)
        rhs.append(f)
    return d, spaces, mats, rhs


def test_f_cycle_matches_oracle(ctx):
    """One F(2,2) cycle (CycleControl F, src/Multigrid/CycleControl.C:17-111)
    on three levels against the oracle's cycle, and FGMRES with it: the same
    iteration count within one and the solution to 1e-7."""
    nref = 3
    kw = dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2)
    d, spaces, mats, rhs = three_level_cd3d(ctx, nref, kw)
    osp = [O.build_space(2, nref, nref - 2 + l, 2)[0] for l in range(3)]
    vals = [A.get_entries() for A in mats]
    prm = dict(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
               residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=200,
               multigrid_cycle_type="F", multigrid_smoother="jacobi", multigrid_smoother_coarse="jacobi",
               multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2, jacobi_damp=0.5,
               multigrid_coarse_residual=0.1, multigrid_coarse_max_n_iterations=10, renumber_levels=0)
    sol = pm.Solver(ctx, pm.solver_params(**prm))
    sol.update_matrix(mats, spaces)
    v = rhs[-1].clone()
    z = torch.zeros_like(v)
    sol.apply_preconditioner(v, z)
    op = O.solver_params(iterative_solver_type=0, preconditioner=2, gmres_restart=30, residual_tolerance=0.0,
                         residual_reduction=1e-8, max_n_iterations=200, mg_cycle=2, mg_smoother_coarse=0,
                         mg_n_pre_smooth=2, mg_n_post_smooth=2, jacobi_damp=0.5, mg_coarse_residual=0.1,
                         mg_coarse_max_n_iterations=10)
    zo = O.mg_apply(op, osp, vals, v.cpu().numpy())
    assert np.max(np.abs(z.cpu().numpy() - zo)) <= 1e-10 * np.max(np.abs(zo))
    # an F cycle is not a V cycle: the planted cycle type must differ
    op_v = O.solver_params(preconditioner=2, mg_cycle=0, mg_smoother_coarse=0, mg_n_pre_smooth=2,
                           mg_n_post_smooth=2, jacobi_damp=0.5, mg_coarse_residual=0.1,
                           mg_coarse_max_n_iterations=10)
    zv = O.mg_apply(op_v, osp, vals, v.cpu().numpy())
    assert np.max(np.abs(zv - zo)) > 1e-6 * np.max(np.abs(zo))
    fs = spaces[-1]
    b = rhs[-1]
    x = torch.zeros_like(b)
    x[fs.active_bound:] = b[fs.active_bound:]
    rep = sol.solve(b, x)
    bo = b.cpu().numpy()
    x0 = np.zeros_like(bo)
    x0[fs.active_bound:] = bo[fs.active_bound:]
    xo, it, hist, st = O.solve(op, osp, vals, bo, x0)
    assert rep.converged and abs(rep.iterations - it) <= 1, (rep.iterations, it)
    assert np.max(np.abs(x.cpu().numpy() - xo)) <= 1e-7 * np.max(np.abs(xo))


def test_cd2d_poisson_64_cg_jacobi_config(ctx):
    """CD2D configuration: -lap u = f on the unit square, 64 x 64 Q1
    (4,225 DOFs), u = sin(pi x) sin(pi y) with Dirichlet data on the whole
    boundary, Gauss 3x3, PCG + Jacobi (Iteration_cg.C:22-79 with
    Iteration_jacobi as preconditioner) to a relative residual of 1e-10.
    Against the oracle: CSR bit-equal, entries 1e-12, the same iteration count
    within one, the solution to 1e-10 relative, L2/H1 errors to 1e-8 relative;
    against the mathematics: the discrete solution converges at second order
    in L2 (the 32 x 32 error is 4x the 64 x 64 one within 3 %)."""
    errs = {}
    for nref in (5, 6):
        prm = pm.cd_params("sine_2d", eps=1.0)
        sp = pm.solver_params(iterative_solver_type="cg", preconditioner="jacobi", residual_tolerance=1e-10,
                              residual_reduction=0.0, max_n_iterations=5000)
        cd = pm.ConvectionDiffusion(ctx, pm.TDomain("UnitSquare", nref), 1, prm, sp)
        cd.assemble()
        rep = cd.solve()
        err = np.array(cd.output())
        errs[nref] = err
        s, A = cd.spaces[-1], cd.matrices[-1]
        osp, _ = O.build_space(0, nref, nref, 1)
        op = O.cd_params(example=6, eps=1.0)
        val, rhs = O.assemble(osp, op)
        rp, col = A.get_structure()
        assert np.array_equal(rp, osp.rp) and np.array_equal(col, osp.col)
        assert np.max(np.abs(A.get_entries() - val)) <= 1e-12 * np.max(np.abs(val))
        x0 = np.zeros(osp.n_dof)
        x0[osp.active:] = rhs[osp.active:]
        params = O.solver_params(iterative_solver_type=1, preconditioner=1, residual_tolerance=1e-10,
                                 residual_reduction=0.0, max_n_iterations=5000)
        xo, it, hist, st = O.solve(params, [osp], [val], rhs, x0)
        print("cd2d", nref, "gpu", rep.iterations, "oracle", it, "errors", err, flush=True)
        assert rep.converged and abs(rep.iterations - it) <= 1, (rep.iterations, it)
        xg = cd.solution.cpu().numpy()
        assert np.max(np.abs(xg - xo)) <= 1e-10 * np.max(np.abs(xo))
        r = rhs.copy()
        O.apply_scaled_add(osp.n_dof, osp.active, osp.rp, osp.col, val, xg, r, -1.0)
        assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(rhs)
        eo = O.errors(osp, op, xo)
        assert np.all(np.abs(err[:2] - eo[:2]) <= 1e-8 * eo[:2]), (err, eo)
    ratio = errs[5][0] / errs[6][0]
    assert abs(ratio - 4.0) < 0.12, ratio
    assert abs(errs[5][1] / errs[6][1] - 2.0) < 0.06


def test_saddle_velocity_block_sell_path(ctx):
    """The three-component SELL pass of the velocity block (one lane per row),
    forced on a small system, against the three CSR launches and the oracle's
    block operator (1e-12), and an FGMRES + Vanka solve with it."""
    nref = 2
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    g = pm.NSESystem(ctx, d, nref)
    o = O.NseOracle(nref, nref)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(3 * o.n_u)
    rhs = g.zeros()
    g.assemble("u2_p1", 1e-2, dev(w), rhs)
    A, B, r = o.assemble(1e-2, w)
    x = rng.standard_normal(o.n)
    yo = o.apply(A, B, x)
    out = {}
    for rows in (131072, 0):
        g.set_velocity_sell_rows(rows)
        y = g.zeros()
        g.apply(dev(x), y)
        out[rows] = y.cpu().numpy()
        assert np.max(np.abs(out[rows] - yo)) <= 1e-12 * np.max(np.abs(yo)), rows
    assert np.max(np.abs(out[0] - out[131072])) <= 1e-13 * np.max(np.abs(yo))
    g.set_velocity_sell_rows(131072)


def _exp_sin_rhs(X, eps, b, c):
    """-eps lap u + b.grad u + c u of u = exp(-x) sin(pi y) sin(pi z)"""
    x, y, z = X[..., 0], X[..., 1], X[..., 2]
    u = np.exp(-x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    ux = -u
    uy = np.pi * np.exp(-x) * np.cos(np.pi * y) * np.sin(np.pi * z)
    uz = np.pi * np.exp(-x) * np.sin(np.pi * y) * np.cos(np.pi * z)
    lap = (1 - 2 * np.pi ** 2) * u
    return -eps * lap + b[0] * ux + b[1] * uy + b[2] * uz + c * u


def _variable_table(X):
    """coefficients of a CoeffFct that no built-in example has: variable
    diffusion, rotating convection, variable reaction and source"""
    x, y, z = X[..., 0], X[..., 1], X[..., 2]
    T = np.zeros(X.shape[:2] + (8,))
    T[..., 0] = 1e-2 * (1 + x * y)
    T[..., 1] = 1 + y
    T[..., 2] = 0.25 * x - z
    T[..., 3] = 0.5 * x
    T[..., 4] = 1 + x * y * z
    T[..., 5] = np.sin(3 * x) + y * z + y
    return T


@pytest.mark.parametrize("rule", [0, 2])
def test_coefficient_table_reproduces_the_example(ctx, rule):
    """pmx_cd_params.dev_coeff_table, the route of a user CoeffFct3D
    (LocalAssembling::GetParameters, LocalAssembling.C:766-840): the
    exp-sin example's coefficients sampled at pmx_assembly_quad_points give
    the example's SUPG assembly on the same (quadrature-kernel) route -- the
    matrix bit for bit (the same constants), the load vector within 1e-13 (f
    evaluated by numpy instead of the kernel) -- and agree with the
    element-GEMM route the constant coefficients take by default to 1e-13."""
    nref = 2
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    s = pm.TFESpace(d, nref, 2, ctx)
    eps, b, c = 1e-2, (0.3, -0.2, 0.5), 0.4
    kw = dict(disc="supg", eps=eps, b=b, c=c, hexa_quad_rule=rule)
    A0 = pm.FEMatrix(ctx, s)
    f0 = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    pm.assemble_cd(ctx, s, pm.cd_params("exp_sin_3d", assembly_variant=1, **kw), A0, f0)
    Ag = pm.FEMatrix(ctx, s)
    fg = torch.zeros_like(f0)
    pm.assemble_cd(ctx, s, pm.cd_params("exp_sin_3d", **kw), Ag, fg)
    X = s.quad_points(rule)
    assert X.shape == (s.n_cells, 64 if rule == 0 else 27, 3)
    T = np.zeros(X.shape[:2] + (8,))
    T[..., 0], T[..., 1:4], T[..., 4] = eps, b, c
    T[..., 5] = _exp_sin_rhs(X, eps, b, c)
    A1 = pm.FEMatrix(ctx, s)
    f1 = torch.zeros_like(f0)
    pm.assemble_cd(ctx, s, pm.cd_params("exp_sin_3d", coeff_table=dev(T), **kw), A1, f1)
    assert np.array_equal(A1.get_entries(), A0.get_entries())
    fa, fb = f0.cpu().numpy(), f1.cpu().numpy()
    assert np.max(np.abs(fa - fb)) <= 1e-13 * np.max(np.abs(fa))
    ga, gb = A0.get_entries(), Ag.get_entries()
    assert np.max(np.abs(ga - gb)) <= 1e-13 * np.max(np.abs(ga))
    assert np.max(np.abs(fa - fg.cpu().numpy())) <= 1e-13 * np.max(np.abs(fa))


@pytest.mark.parametrize("geo,nref,order,disc", [(2, 2, 2, "galerkin"), (2, 2, 2, "supg"), (2, 2, 1, "supg"),
                                                 (0, 3, 2, "supg"), (0, 3, 1, "galerkin")])
def test_coefficient_table_variable_coefficients_match_oracle(ctx, geo, nref, order, disc):
    """A coefficient function no example has (variable eps, b, c, f) through
    dev_coeff_table against the oracle fed the same function at its own
    quadrature points (which are the product's, in the same order): entries
    and load vector within 1e-12; SUPG takes its delta from each point's
    coefficients, as BilinearAssemble_SD does."""
    rule = 2
    d = pm.TDomain({0: "UnitSquare", 2: "Default_UnitCube_Hexa"}[geo], nref)
    s = pm.TFESpace(d, nref, order, ctx)
    Xg = s.quad_points(rule)
    osp, _ = O.build_space(geo, nref, nref, order)
    Xo = O.quad_points(osp, rule)
    assert Xg.shape == Xo.shape and np.max(np.abs(Xg - Xo)) <= 1e-15
    ex = "exp_sin_3d" if geo == 2 else "sine_2d"
    A = pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    pm.assemble_cd(ctx, s, pm.cd_params(ex, disc=disc, hexa_quad_rule=rule, coeff_table=dev(_variable_table(Xg))),
                   A, f)
    val, rhs = O.assemble(osp, O.cd_params(example=5 if geo == 2 else 6, disc=1 if disc == "supg" else 0,
                                           hexa_rule=rule, table=_variable_table(Xo)))
    rp, col = A.get_structure()
    assert np.array_equal(rp, osp.rp) and np.array_equal(col, osp.col)
    assert np.max(np.abs(A.get_entries() - val)) <= 1e-12 * np.max(np.abs(val))
    assert np.max(np.abs(f.cpu().numpy() - rhs)) <= 1e-12 * np.max(np.abs(rhs))


def test_host_dirichlet_values(ctx):
    """pmx_cd_params.host_dirichlet: boundary values computed by the caller
    at pmx_space_get_dof_coords replace the example's (Assemble3D.C:593-667
    sets the Dirichlet rhs entries to the boundary values): the Dirichlet
    part of the rhs is the given array, the active part and the matrix are
    unchanged; a too short array is refused before any kernel."""
    nref = 2
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    s = pm.TFESpace(d, nref, 2, ctx)
    ab = s.active_bound
    xyz = s.dof_coords()[ab:]
    g = 1.0 + xyz[:, 0] ** 2 - 2.0 * xyz[:, 1] * xyz[:, 2]
    kw = dict(disc="supg", eps=1e-2, b=(0.3, -0.2, 0.5), hexa_quad_rule=2)
    A0, A1 = pm.FEMatrix(ctx, s), pm.FEMatrix(ctx, s)
    f0 = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    f1 = torch.zeros_like(f0)
    pm.assemble_cd(ctx, s, pm.cd_params("exp_sin_3d", **kw), A0, f0)
    pm.assemble_cd(ctx, s, pm.cd_params("exp_sin_3d", dirichlet=g, **kw), A1, f1)
    assert np.array_equal(A0.get_entries(), A1.get_entries())
    a, b = f0.cpu().numpy(), f1.cpu().numpy()
    assert np.array_equal(a[:ab], b[:ab])
    assert np.array_equal(b[ab:], g)
    # the example's own boundary values at the same points, as a cross-check
    u = np.exp(-xyz[:, 0]) * np.sin(np.pi * xyz[:, 1]) * np.sin(np.pi * xyz[:, 2])
    assert np.max(np.abs(a[ab:] - u)) <= 1e-15
    with pytest.raises(ValueError):
        pm.assemble_cd(ctx, s, pm.cd_params("exp_sin_3d", dirichlet=g[:-1], **kw), A1, f1)


def test_objects_outlive_their_context(ctx):
    """pmx_ctx_destroy before the objects created on the context (a garbage
    collector may finalise them in any order): each object holds a reference
    to the context, so its destroy still finds the stream; the context is
    freed with the last of them."""
    import gc
    c2 = pm.Context(0)
    d = pm.TDomain("Default_UnitCube_Hexa", 1)
    s = pm.TFESpace(d, 1, 2, c2)
    A = pm.FEMatrix(c2, s)
    sol = pm.Solver(c2, pm.solver_params(iterative_solver_type="fgmres", preconditioner="jacobi"))
    g = pm.NSESystem(c2, d, 1)
    pm._destroy("pmx_ctx_destroy", c2.h)
    c2.h = None
    del sol, A, g, s
    gc.collect()
    # the library still works on the other context
    x = torch.ones(8, dtype=torch.float64, device="cuda")
    assert abs(pm.vector_norm(ctx, x) - 8 ** 0.5) < 1e-14
