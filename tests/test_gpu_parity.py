"""Parity of the HIP path (through the C ABI) with the CPU oracle and the
reference's golden values.  FP64 tolerances: assembled entries and SpMV results
within 1e-12 relative to the row scale; golden errors within the reference's
own tolerances."""
import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from mymoonmd_amd._lib import PmxError
from conftest import iterative_7x7
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEO = {"UnitSquare": 0, "TwoTriangles": 1, "Default_UnitCube_Hexa": 2}
EX = {"const": 0, "sine_laplace_2d": 1, "p0_3d": 2, "p1_3d": 3, "p2_3d": 4, "exp_sin_3d": 5, "sine_2d": 6}
B_CD3D = (1 / 14 ** 0.5, 2 / 14 ** 0.5, 3 / 14 ** 0.5)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def gpu_assemble(ctx, geo, nref, order, bc, params):
    d = pm.TDomain(geo, nref)
    s = pm.TFESpace(d, nref, order, ctx, bc=bc)
    A = pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    pm.assemble_cd(ctx, s, params, A, f)
    return d, s, A, f


def oracle_assemble(geo, nref, order, bc, **kw):
    cond2d = (lambda c, t: 1 if c == 1 else 0) if bc == "sine_laplace" else None
    sp, info = O.build_space(GEO[geo], nref, nref, order, cond2d=cond2d)
    p = O.cd_params(sine_laplace_bc=(bc == "sine_laplace"), **kw)
    val, rhs = O.assemble(sp, p)
    return sp, p, val, rhs


ASM_CASES = [
    ("UnitSquare", 3, 1, "sine_laplace", dict(example="sine_laplace_2d"), {}),
    ("UnitSquare", 3, 2, "sine_laplace", dict(example="sine_laplace_2d"), {}),
    ("TwoTriangles", 3, 1, "all_dirichlet", dict(example="sine_2d", eps=0.01, b=(1.0, 0.5), c=0.5), {}),
    ("TwoTriangles", 3, 1, "all_dirichlet", dict(example="sine_2d", disc="supg", eps=0.001, b=(1.0, 0.5)), {}),
    ("UnitSquare", 3, 1, "all_dirichlet", dict(example="sine_2d", disc="supg", eps=0.001, b=(1.0, 0.5)), {}),
    ("Default_UnitCube_Hexa", 2, 1, "all_dirichlet", dict(example="p1_3d", eps=1.0, b=(1, 2, 3)), {}),
    ("Default_UnitCube_Hexa", 2, 2, "all_dirichlet", dict(example="p2_3d", eps=1.0, b=(4, 3, 2), c=1.0), {}),
    ("Default_UnitCube_Hexa", 3, 2, "all_dirichlet",
     dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2), {}),
    ("Default_UnitCube_Hexa", 2, 2, "all_dirichlet",
     dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=0), {}),
    ("TwoTriangles", 3, 1, "all_dirichlet", dict(example="sine_2d", disc="gls", eps=0.001, b=(1.0, 0.5), c=0.5), {}),
    ("UnitSquare", 3, 2, "all_dirichlet", dict(example="sine_2d", disc="gls", eps=0.01, b=(1.0, 0.5), c=0.5), {}),
    ("Default_UnitCube_Hexa", 3, 2, "all_dirichlet",
     dict(example="exp_sin_3d", disc="gls", eps=1e-3, b=B_CD3D, c=0.5, hexa_quad_rule=2), {}),
    ("Default_UnitCube_Hexa", 2, 2, "all_dirichlet",
     dict(example="exp_sin_3d", disc="gls", eps=1e-2, b=B_CD3D, hexa_quad_rule=0), {}),
]


def oracle_kw(kw):
    out = dict(example=EX[kw.get("example", "const")], disc={"galerkin": 0, "supg": 1, "gls": 2}[kw.get("disc", "galerkin")],
               eps=kw.get("eps", 1.0), c=kw.get("c", 0.0), hexa_rule=kw.get("hexa_quad_rule", 0))
    b = list(kw.get("b", (0, 0, 0))) + [0.0] * 3
    out["b"] = b[:3]
    return out


@pytest.mark.parametrize("geo,nref,order,bc,kw,_", ASM_CASES)
def test_assembly_matches_oracle(ctx, geo, nref, order, bc, kw, _):
    print("case", geo, nref, order, kw, flush=True)
    d, s, A, f = gpu_assemble(ctx, geo, nref, order, bc, pm.cd_params(**kw))
    sp, p, val, rhs = oracle_assemble(geo, nref, order, bc, **oracle_kw(kw))
    rp, col = A.get_structure()
    assert np.array_equal(rp, sp.rp) and np.array_equal(col, sp.col)
    gv = A.get_entries()
    scale = np.max(np.abs(val))
    assert np.max(np.abs(gv - val)) <= 1e-12 * scale
    gr = f.cpu().numpy()
    assert np.max(np.abs(gr - rhs)) <= 1e-12 * max(1.0, np.max(np.abs(rhs)))


def test_spmv_matches_oracle(ctx):
    kw = dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2)
    d, s, A, f = gpu_assemble(ctx, "Default_UnitCube_Hexa", 4, 2, "all_dirichlet", pm.cd_params(**kw))
    rp, col = A.get_structure()
    val = A.get_entries()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(s.n_dof)
    y0 = rng.standard_normal(s.n_dof)
    for lanes in (0, 8, 16, 32, 64):   # 0: SELL-64 kernel
        A.set_spmv_lanes(lanes)
        yg = dev(y0)
        A.apply_scaled_add(dev(x), yg, -0.7)
        yo = y0.copy()
        O.apply_scaled_add(s.n_dof, s.active_bound, rp, col, val, x, yo, -0.7)
        absA = np.abs(val)
        bound = np.zeros(s.n_dof)
        O.apply_scaled_add(s.n_dof, s.active_bound, rp, col, absA, np.abs(x), bound, 0.7)
        assert np.all(np.abs(yg.cpu().numpy() - yo) <= 1e-13 * (bound + np.abs(y0)) + 1e-300), lanes
    A.set_spmv_lanes(0)
    yg = torch.empty_like(dev(x))
    A.apply(dev(x), yg)
    yo = np.zeros(s.n_dof)
    O.apply_scaled_add(s.n_dof, s.active_bound, rp, col, val, x, yo, 1.0)
    assert np.allclose(yg.cpu().numpy(), yo, rtol=1e-12, atol=1e-12 * np.max(np.abs(yo)))


def test_spmv_matrix_test_golden(ctx, golden):
    g = golden["matrix_test"]

    def dense(m):
        a = np.zeros((m["n_rows"], m["n_cols"]))
        for r in range(m["n_rows"]):
            for k in range(m["row_ptr"][r], m["row_ptr"][r + 1]):
                a[r, m["col"][k]] = m["val"][k]
        return a

    CB = dense(g["C"]) @ dense(g["B"])
    rp, col, val = [0], [], []
    for r in range(4):
        nz = np.nonzero(CB[r])[0]
        col += list(nz)
        val += list(CB[r, nz])
        rp.append(len(col))
    M = pm.FEMatrix(ctx, csr=(rp, col, val, 4))
    y = torch.zeros(4, dtype=torch.float64, device="cuda")
    M.apply_scaled_add(dev(g["x"]), y, -0.5)
    ref = np.array(g["y_minus_half_CB_x"])
    assert np.all(np.abs((y.cpu().numpy() - ref) / ref) < 1e-14)


@pytest.mark.parametrize("method,prec", [("jacobi", "no_preconditioner"), ("richardson", "jacobi"),
                                         ("cg", "jacobi"), ("fgmres", "jacobi"),
                                         ("fgmres", "no_preconditioner")])
def test_iterative_solvers_7x7(ctx, golden, method, prec):
    g = golden["iterative_solver_test"]
    Mx, rp, col, val = iterative_7x7(g)
    M = pm.FEMatrix(ctx, csr=(rp, col, val, 7))
    exact = np.arange(7, dtype=float)
    b = dev(Mx @ exact)
    x = torch.zeros(7, dtype=torch.float64, device="cuda")
    sol = pm.Solver(ctx, pm.solver_params(iterative_solver_type=method, preconditioner=prec,
                                          residual_tolerance=g["residual_tolerance"],
                                          max_n_iterations=g["max_n_iterations"]))
    sol.update_matrix(M)
    rep = sol.solve(b, x)
    assert np.linalg.norm(x.cpu().numpy() - exact) < 100 * g["residual_tolerance"]
    # same iteration count as the oracle restatement
    params = O.solver_params(iterative_solver_type=pm.IT_SOLVER[method], preconditioner=pm.PRECOND[prec],
                             residual_tolerance=g["residual_tolerance"], max_n_iterations=g["max_n_iterations"])
    xo, it, hist, st = O.solve(params, [O.csr_general(7, 7, rp, col)], [val], Mx @ exact, np.zeros(7))
    assert abs(rep.iterations - it) <= 1


@pytest.mark.parametrize("geo,order", [("UnitSquare", 1), ("UnitSquare", 2), ("TwoTriangles", 1)])
def test_cd2d_golden_errors_on_gpu(ctx, golden, geo, order):
    ref = np.array(golden["cd2d_test"][geo][str(order)])
    prm = pm.cd_params("sine_laplace_2d", eps=1.0)
    sp = pm.solver_params(iterative_solver_type="fgmres", preconditioner="no_preconditioner",
                          residual_tolerance=1e-14, max_n_iterations=3000, gmres_restart=400)
    cd = pm.ConvectionDiffusion(ctx, pm.TDomain(geo, 2), order, prm, sp, bc="sine_laplace")
    cd.assemble()
    cd.solve()
    err = np.array(cd.output())
    assert np.all(np.abs(err - ref) < golden["cd2d_test"]["tolerance"]), (err, ref)


def test_cd2d_golden_errors_fgmres_multigrid(ctx, golden):
    """cd2d_test's third set-up: FGMRES + multigrid (W, 2 levels, jacobi 2/2, direct coarse)."""
    ref = np.array(golden["cd2d_test"]["UnitSquare"]["1"])
    prm = pm.cd_params("sine_laplace_2d", eps=1.0)
    sp = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid",
                          residual_tolerance=1e-13, max_n_iterations=1000, multigrid_cycle_type="W",
                          multigrid_smoother="jacobi", multigrid_smoother_coarse="direct_solve",
                          multigrid_correction_damp_factor=1.0, multigrid_n_pre_smooth=2,
                          multigrid_n_post_smooth=2)
    cd = pm.ConvectionDiffusion(ctx, pm.TDomain("UnitSquare", 2), 1, prm, sp, n_mg_levels=2,
                                bc="sine_laplace")
    cd.assemble()
    cd.solve()
    err = np.array(cd.output())
    assert np.all(np.abs(err - ref) < golden["cd2d_test"]["tolerance"]), (err, ref)


@pytest.mark.parametrize("idx", [0, 1, 2])
@pytest.mark.parametrize("solver", ["multigrid", "jacobi"])
@pytest.mark.parametrize("disc", ["galerkin", "gls"])
def test_cd3d_zero_error_examples_on_gpu(ctx, golden, idx, solver, disc):
    name, order, eps, b, c = golden["cd3d_solvers_test"]["examples"][idx]
    prm = pm.cd_params(name, disc=disc, eps=eps, b=b, c=c)
    if disc == "gls" and solver == "jacobi" and order == 2:
        # Richardson + Jacobi cannot converge on the Q2 GLS matrix: the
        # spectral radius of I - D^-1 A on the active block is 1.53 (the
        # multigrid variant covers this case); the solver must report the
        # divergence (IterativeMethod.h:105-131) rather than return
        sp = pm.solver_params(iterative_solver_type="richardson", preconditioner="jacobi",
                              residual_tolerance=1e-10, residual_reduction=0.0, max_n_iterations=1000,
                              divergence_factor=1e3)
        cd = pm.ConvectionDiffusion(ctx, pm.TDomain("Default_UnitCube_Hexa", 2), order, prm, sp)
        cd.assemble()
        A, s = cd.matrices[-1], cd.spaces[-1]
        rp, col = A.get_structure()
        v = A.get_entries()
        ab = s.active_bound
        D = np.zeros((s.n_dof, s.n_dof))
        for r in range(s.n_dof):
            D[r, col[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
        Aa = D[:ab, :ab]
        rho = np.max(np.abs(np.linalg.eigvals(np.eye(ab) - Aa / np.diag(Aa)[:, None])))
        assert rho > 1.5, rho
        with pytest.raises(PmxError) as e:
            cd.solve()
        assert e.value.code == 4
        return
    if solver == "multigrid":
        # GLS adds delta (L phi_j, L phi_i) with -eps lap phi (eps = 1): the
        # undamped Jacobi smoother of this set-up contracts more slowly there
        sp = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid",
                              residual_tolerance=1e-15, max_n_iterations=100 if disc == "galerkin" else 600,
                              multigrid_cycle_type="W",
                              multigrid_smoother="jacobi", multigrid_smoother_coarse="direct_solve",
                              multigrid_correction_damp_factor=0.8, multigrid_n_pre_smooth=3,
                              multigrid_n_post_smooth=3)
        tol, nl = golden["cd3d_solvers_test"]["tolerance_multigrid"], 2
    else:
        sp = pm.solver_params(iterative_solver_type="richardson", preconditioner="jacobi",
                              residual_tolerance=1e-10, residual_reduction=0.0, max_n_iterations=1000,
                              min_n_iterations=5)
        tol, nl = golden["cd3d_solvers_test"]["tolerance_jacobi"], 1
    cd = pm.ConvectionDiffusion(ctx, pm.TDomain("Default_UnitCube_Hexa", 2), order, prm, sp,
                                n_mg_levels=nl)
    cd.assemble()
    cd.solve()
    err = cd.output()
    assert err[0] < tol and err[1] < tol, err


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


def oracle_levels(spaces, mats, nref):
    osp = []
    for l, s in enumerate(spaces):
        sp, _ = O.build_space(2, nref, nref - 2 + l, 2)
        osp.append(sp)
    vals = [A.get_entries() for A in mats]
    return osp, vals


MG_KW = dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2)


def test_multigrid_cycle_matches_oracle(ctx):
    """One V(2,2) cycle, damped Jacobi 0.8, coarse Jacobi iterations (3 levels)."""
    nref = 3
    d, spaces, mats, rhs = three_level_cd3d(ctx, nref, MG_KW)
    osp, vals = oracle_levels(spaces, mats, nref)
    for coarse in ("jacobi", "direct_solve"):
        prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid",
                               multigrid_cycle_type="V", multigrid_smoother="jacobi",
                               multigrid_smoother_coarse=coarse, multigrid_n_pre_smooth=2,
                               multigrid_n_post_smooth=2, jacobi_damp=0.8,
                               multigrid_correction_damp_factor=1.0, multigrid_coarse_residual=0.1,
                               multigrid_coarse_max_n_iterations=10)
        sol = pm.Solver(ctx, prm)
        sol.update_matrix(mats, spaces)
        v = rhs[-1].clone()
        z = torch.zeros_like(v)
        sol.apply_preconditioner(v, z)
        op = O.solver_params(preconditioner=2, mg_cycle=0, mg_smoother_coarse=1 if coarse == "direct_solve" else 0,
                             mg_n_pre_smooth=2, mg_n_post_smooth=2, jacobi_damp=0.8,
                             mg_correction_damp=1.0, mg_coarse_residual=0.1, mg_coarse_max_n_iterations=10)
        zo = O.mg_apply(op, osp, vals, v.cpu().numpy())
        zg = z.cpu().numpy()
        assert np.max(np.abs(zg - zo)) <= 1e-10 * np.max(np.abs(zo)), coarse


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("cycle", ["V", "W"])
def test_zero_start_sweep_and_renumbering(ctx, precision, cycle):
    """Two execution choices of the multigrid.  The first Jacobi step on a
    level whose solution was reset multiplies only the Dirichlet-column
    entries: cycles and solves agree bit for bit with the full sweep (the
    preconditioned vector has non-zero Dirichlet entries, so the Dirichlet
    couplings are exercised).  The levels above the coarsest run in a
    lexicographic DOF order with row entries sorted by the new column: only
    the rounding of the row sums changes (1e-12 FP64, 1e-5 with FP32/bf16),
    and the solve takes the same number of iterations; storing those levels'
    columns as 16-bit offsets (renumber_levels 2) changes nothing; folding
    the Dirichlet-column products into per-reset sums (renumber_levels 3)
    sums them first instead of last (rounding only); packing a row's slots by
    four (renumber_levels 4) changes nothing; nor does grouping the rows of a
    slice by their offset pattern (renumber_levels 5, and 6 with packed
    slots)."""
    nref = 3
    d, spaces, mats, rhs = three_level_cd3d(ctx, nref, MG_KW)
    fs = spaces[-1]
    out = {}
    for shortcut, renumber in ((0, 0), (1, 0), (0, 1), (1, 1), (0, 2), (1, 2), (0, 3), (1, 3), (0, 4), (1, 4),
                               (0, 5), (1, 5), (0, 6), (1, 6)):
        for coarse in ("jacobi", "direct_solve"):
            prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
                                   residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=100,
                                   multigrid_cycle_type=cycle, multigrid_smoother="jacobi",
                                   multigrid_smoother_coarse=coarse, multigrid_n_pre_smooth=2,
                                   multigrid_n_post_smooth=2, jacobi_damp=0.5, precision=precision,
                                   zero_start_shortcut=shortcut, renumber_levels=renumber, refinement=0)
            sol = pm.Solver(ctx, prm)
            sol.update_matrix(mats, spaces)
            z = torch.zeros_like(rhs[-1])
            sol.apply_preconditioner(rhs[-1].clone(), z)
            x = torch.zeros_like(rhs[-1])
            x[fs.active_bound:] = rhs[-1][fs.active_bound:]
            rep = sol.solve(rhs[-1], x)
            out[(shortcut, renumber, coarse)] = (z.cpu().numpy(), x.cpu().numpy(), rep.iterations)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for coarse in ("jacobi", "direct_solve"):
        for renumber in (0, 1, 2, 3, 4, 5, 6):
            a, b = out[(0, renumber, coarse)], out[(1, renumber, coarse)]
            assert np.array_equal(a[0], b[0]), (renumber, coarse)
            assert np.array_equal(a[1], b[1]) and a[2] == b[2], (renumber, coarse)
        for r0, r1 in ((1, 2), (3, 4), (3, 5), (3, 6)):
            a, b = out[(1, r0, coarse)], out[(1, r1, coarse)]
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2], (coarse, r0, r1)
        for r0, r1 in ((0, 1), (1, 3)):
            a, b = out[(1, r0, coarse)], out[(1, r1, coarse)]
            assert np.max(np.abs(a[0] - b[0])) <= tol * np.max(np.abs(a[0])), (coarse, r0, r1)
            assert abs(a[2] - b[2]) <= 1 and np.max(np.abs(a[1] - b[1])) <= 1e-7 * np.max(np.abs(a[1])), (coarse, r0, r1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_transfer_sell_matches_csr(ctx, precision):
    """Prolongation and restriction from the SELL copy of P and P^T (one lane
    per row, sequential row sums) against the CSR kernels (1-8 lanes per row):
    one V-cycle agrees to rounding (1e-12 FP64, 1e-5 FP32), and FGMRES+MG
    takes the same number of iterations to the same solution."""
    nref = 3
    d, spaces, mats, rhs = three_level_cd3d(ctx, nref, MG_KW)
    fs = spaces[-1]
    out = {}
    for ts in (0, 2):
        prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
                               residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=100,
                               multigrid_cycle_type="V", multigrid_smoother="jacobi",
                               multigrid_smoother_coarse="direct_solve", multigrid_n_pre_smooth=2,
                               multigrid_n_post_smooth=2, jacobi_damp=0.5, precision=precision,
                               transfer_sell=ts)
        sol = pm.Solver(ctx, prm)
        sol.update_matrix(mats, spaces)
        z = torch.zeros_like(rhs[-1])
        sol.apply_preconditioner(rhs[-1].clone(), z)
        x = torch.zeros_like(rhs[-1])
        x[fs.active_bound:] = rhs[-1][fs.active_bound:]
        rep = sol.solve(rhs[-1], x)
        out[ts] = (z.cpu().numpy(), x.cpu().numpy(), rep.iterations)
    tol = 1e-12 if precision == "fp64" else 1e-5
    a, b = out[0], out[2]
    assert np.max(np.abs(a[0] - b[0])) <= tol * np.max(np.abs(a[0]))
    assert abs(a[2] - b[2]) <= 1 and np.max(np.abs(a[1] - b[1])) <= 1e-7 * np.max(np.abs(a[1]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pattern_transfers_match_sell(ctx, precision):
    """Transfers from pattern slices (rows sharing column offsets and weights,
    Dirichlet columns left out; transfer_sell 4 puts every renumbered level
    on them) against the SELL copy of P and P^T (transfer_sell 2): the row
    sums run in the same column order and the left-out columns only meet
    zeros, so one V-cycle and the whole FGMRES+MG solve are the same bit for
    bit."""
    nref = 3
    d, spaces, mats, rhs = three_level_cd3d(ctx, nref, MG_KW)
    fs = spaces[-1]
    out = {}
    for ts in (2, 4):
        prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
                               residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=100,
                               multigrid_cycle_type="V", multigrid_smoother="jacobi",
                               multigrid_smoother_coarse="direct_solve", multigrid_n_pre_smooth=2,
                               multigrid_n_post_smooth=2, jacobi_damp=0.5, precision=precision,
                               transfer_sell=ts)
        sol = pm.Solver(ctx, prm)
        sol.update_matrix(mats, spaces)
        z = torch.zeros_like(rhs[-1])
        sol.apply_preconditioner(rhs[-1].clone(), z)
        x = torch.zeros_like(rhs[-1])
        x[fs.active_bound:] = rhs[-1][fs.active_bound:]
        rep = sol.solve(rhs[-1], x)
        out[ts] = (z.cpu().numpy(), x.cpu().numpy(), rep.iterations)
    a, b = out[2], out[4]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_fgmres_multigrid_solution_matches_oracle(ctx):
    """FGMRES(30) + MG V(2,2) to 1e-8.  Damping 0.5 keeps the Jacobi smoother
    contractive for this Q2 SUPG operator (lambda_max(D^-1 A) ~ 3), so the
    iteration count is insensitive to round-off and comparable across paths."""
    nref = 3
    d, spaces, mats, rhs = three_level_cd3d(ctx, nref, MG_KW)
    osp, vals = oracle_levels(spaces, mats, nref)
    fs = spaces[-1]
    b = rhs[-1]
    x = torch.zeros_like(b)
    x[fs.active_bound:] = b[fs.active_bound:]
    kw = dict(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
              residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=200,
              multigrid_cycle_type="V", multigrid_smoother="jacobi", multigrid_smoother_coarse="jacobi",
              multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2, jacobi_damp=0.5)
    sol = pm.Solver(ctx, pm.solver_params(**kw))
    sol.update_matrix(mats, spaces)
    rep = sol.solve(b, x)
    op = O.solver_params(iterative_solver_type=0, preconditioner=2, gmres_restart=30, residual_tolerance=0.0,
                         residual_reduction=1e-8, max_n_iterations=200, mg_cycle=0, mg_smoother_coarse=0,
                         mg_n_pre_smooth=2, mg_n_post_smooth=2, jacobi_damp=0.5)
    bo = b.cpu().numpy()
    x0 = np.zeros_like(bo)
    x0[fs.active_bound:] = bo[fs.active_bound:]
    xo, it, hist, st = O.solve(op, osp, vals, bo, x0)
    assert rep.converged and abs(rep.iterations - it) <= 1, (rep.iterations, it)
    xg = x.cpu().numpy()
    assert np.max(np.abs(xg - xo)) <= 1e-7 * np.max(np.abs(xo))
    # FP64 residual of the GPU solution, recomputed by the oracle
    r = bo.copy()
    O.apply_scaled_add(fs.n_dof, fs.active_bound, osp[-1].rp, osp[-1].col, vals[-1], xg, r, -1.0)
    r0 = bo.copy()
    O.apply_scaled_add(fs.n_dof, fs.active_bound, osp[-1].rp, osp[-1].col, vals[-1], x0, r0, -1.0)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(r0)


def test_mixed_precision_reaches_fp64_residual(ctx):
    """FP32 V-cycle on bf16 (row-scaled) matrices inside FP64 FGMRES: the
    FP64-verified relative residual reaches 1e-8 and the solution agrees with
    the FP64 path to the discretisation-irrelevant level (1e-6 relative)."""
    nref = 3
    d, spaces, mats, rhs = three_level_cd3d(ctx, nref, MG_KW)
    fs = spaces[-1]
    b = rhs[-1]
    sols = {}
    for prec in ("fp64", "mixed"):
        x = torch.zeros_like(b)
        x[fs.active_bound:] = b[fs.active_bound:]
        sol = pm.Solver(ctx, pm.solver_params(
            iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
            residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=300,
            multigrid_cycle_type="V", multigrid_smoother="jacobi", multigrid_smoother_coarse="jacobi",
            multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2, jacobi_damp=0.5, precision=prec))
        sol.update_matrix(mats, spaces)
        rep = sol.solve(b, x)
        assert rep.converged
        sols[prec] = x.cpu().numpy()
    val = mats[-1].get_entries()
    rp, col = mats[-1].get_structure()
    bo = b.cpu().numpy()
    x0 = np.zeros_like(bo)
    x0[fs.active_bound:] = bo[fs.active_bound:]
    r0 = bo.copy()
    O.apply_scaled_add(fs.n_dof, fs.active_bound, rp, col, val, x0, r0, -1.0)
    r = bo.copy()
    O.apply_scaled_add(fs.n_dof, fs.active_bound, rp, col, val, sols["mixed"], r, -1.0)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(r0)
    assert np.max(np.abs(sols["mixed"] - sols["fp64"])) <= 1e-6 * np.max(np.abs(sols["fp64"]))


def test_full_size_assembly_properties(ctx):
    """64^3 Q2 (benchmark size): the SUPG+Galerkin operator annihilates constants
    on every active row, Dirichlet rows are identity rows, and the quadratic
    example's nodal interpolant solves the Galerkin system (exact quadrature)."""
    d = pm.TDomain("Default_UnitCube_Hexa", 6)
    s = pm.TFESpace(d, 6, 2, ctx)
    A = pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    pm.assemble_cd(ctx, s, pm.cd_params(**MG_KW), A, f)
    ones = torch.ones(s.n_dof, dtype=torch.float64, device="cuda")
    y = torch.empty_like(ones)
    A.apply(ones, y)
    ab = s.active_bound
    assert torch.max(torch.abs(y[:ab])).item() < 1e-14
    assert torch.all(y[ab:] == 1.0)
    del A, f, y
    A = pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    pm.assemble_cd(ctx, s, pm.cd_params("p2_3d", eps=1.0, b=(4, 3, 2), c=1.0), A, f)
    xyz = s.dof_coords()
    uI = dev(xyz[:, 0] ** 2 + xyz[:, 1] * xyz[:, 2])
    r = f.clone()
    A.apply_scaled_add(uI, r, -1.0)
    assert torch.linalg.norm(r).item() <= 1e-11 * torch.linalg.norm(f).item()
