"""Mixed precision as the north star puts it: FP64 iterative refinement (FP64
residual and solution update with the FP64 matrix) around FP32 FGMRES whose
preconditioner is the FP32 multigrid on bf16 row-scaled matrices and whose
operator is an FP32 (or bf16) copy of the finest matrix.  Checked against the
FP64 GPU path and the oracle: the FP64-verified relative residual reaches the
requested 1e-8 (and 1e-10), the solution agrees with the FP64 path, and the
iteration count stays close to the FP64 FGMRES count."""
import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

B_CD3D = (1 / 14 ** 0.5, 2 / 14 ** 0.5, 3 / 14 ** 0.5)
KW = dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2)


def levels(ctx, nref, nlev):
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    spaces = [pm.TFESpace(d, l, 2, ctx) for l in range(nref - nlev + 1, nref + 1)]
    mats, rhs = [], []
    for s in spaces:
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.assemble_cd(ctx, s, pm.cd_params(**KW), A, f)
        mats.append(A)
        rhs.append(f)
    return d, spaces, mats, rhs


def solve(ctx, spaces, mats, b, precision, reduction, **kw):
    fs = spaces[-1]
    x = torch.zeros_like(b)
    x[fs.active_bound:] = b[fs.active_bound:]
    prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
                           residual_tolerance=0.0, residual_reduction=reduction, max_n_iterations=300,
                           multigrid_cycle_type="V", multigrid_smoother="jacobi",
                           multigrid_smoother_coarse="direct_solve", multigrid_n_pre_smooth=2,
                           multigrid_n_post_smooth=2, jacobi_damp=0.6, precision=precision, **kw)
    sol = pm.Solver(ctx, prm)
    sol.update_matrix(mats, spaces)
    rep = sol.solve(b, x)
    return x, rep, sol


def fp64_residual(A, b, x, x0):
    r = b.clone()
    A.apply_scaled_add(x, r, -1.0)
    r0 = b.clone()
    A.apply_scaled_add(x0, r0, -1.0)
    return torch.linalg.norm(r).item() / torch.linalg.norm(r0).item()


@pytest.mark.parametrize("inner_bf16", [0, 1])
@pytest.mark.parametrize("reduction", [1e-8, 1e-10])
def test_refinement_reaches_the_fp64_residual(ctx, inner_bf16, reduction):
    nref = 4
    print("case", nref, inner_bf16, reduction, flush=True)
    d, spaces, mats, rhs = levels(ctx, nref, 3)
    fs, A, b = spaces[-1], mats[-1], rhs[-1]
    x0 = torch.zeros_like(b)
    x0[fs.active_bound:] = b[fs.active_bound:]
    x64, rep64, _ = solve(ctx, spaces, mats, b, "fp64", reduction)
    # a bf16 inner operator limits each refinement step to a reduction of about
    # rho(A^-1 (A - A_bf16)) (~1e-2 here): its inner solves stop there
    xm, repm, sol = solve(ctx, spaces, mats, b, "mixed", reduction, refinement=1, refinement_inner_bf16=inner_bf16,
                          refinement_inner_reduction=2e-2 if inner_bf16 else 1e-4)
    assert rep64.converged and repm.converged
    rel = fp64_residual(A, b, xm, x0)
    assert rel <= 1.0001 * reduction, rel
    # the solver's own final residual is the FP64 one
    assert abs(repm.final_residual / repm.initial_residual - rel) <= 1e-3 * rel
    if inner_bf16:
        assert repm.iterations <= 2 * rep64.iterations, (repm.iterations, rep64.iterations)
    else:
        assert repm.iterations <= rep64.iterations + 3, (repm.iterations, rep64.iterations)
    diff = torch.max(torch.abs(xm - x64)).item() / torch.max(torch.abs(x64)).item()
    assert diff <= 100 * reduction, diff


def test_refinement_matches_the_oracle_solution(ctx):
    """FP64 refinement to 1e-8 against the oracle's FP64 FGMRES + multigrid to
    1e-10: the solutions agree to the tolerance the residual allows."""
    nref = 3
    d, spaces, mats, rhs = levels(ctx, nref, 2)
    fs, b = spaces[-1], rhs[-1]
    xm, repm, _ = solve(ctx, spaces, mats, b, "mixed", 1e-8, refinement=1)
    osp = [O.build_space(2, nref, nref - 1 + l, 2)[0] for l in range(2)]
    vals = [M.get_entries() for M in mats]
    bo = b.cpu().numpy()
    x0 = np.zeros_like(bo)
    x0[fs.active_bound:] = bo[fs.active_bound:]
    op = O.solver_params(iterative_solver_type=0, preconditioner=2, gmres_restart=30, residual_tolerance=0.0,
                         residual_reduction=1e-10, max_n_iterations=300, mg_cycle=0, mg_smoother_coarse=1,
                         mg_n_pre_smooth=2, mg_n_post_smooth=2, jacobi_damp=0.6)
    xo, it, hist, st = O.solve(op, osp, vals, bo, x0)
    assert np.max(np.abs(xm.cpu().numpy() - xo)) <= 1e-6 * np.max(np.abs(xo))
    # the residual of the refined solution, recomputed by the oracle in FP64
    r = bo.copy()
    O.apply_scaled_add(fs.n_dof, fs.active_bound, osp[-1].rp, osp[-1].col, vals[-1], xm.cpu().numpy(), r, -1.0)
    r0 = bo.copy()
    O.apply_scaled_add(fs.n_dof, fs.active_bound, osp[-1].rp, osp[-1].col, vals[-1], x0, r0, -1.0)
    assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(r0)


def test_refinement_off_is_the_plain_mixed_fgmres(ctx):
    """refinement 0 keeps FGMRES in FP64 (the preconditioner alone in FP32)."""
    d, spaces, mats, rhs = levels(ctx, 3, 3)
    fs, A, b = spaces[-1], mats[-1], rhs[-1]
    x0 = torch.zeros_like(b)
    x0[fs.active_bound:] = b[fs.active_bound:]
    x, rep, _ = solve(ctx, spaces, mats, b, "mixed", 1e-8, refinement=0)
    assert rep.converged and fp64_residual(A, b, x, x0) <= 1.01e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_cycle_replays_equal_the_launched_cycle(ctx, precision):
    """The cycle captured as a HIP graph: three solves in a row (every
    iteration replays the graph), with eager work between them (a
    preconditioner application, a matrix product), give the same bits and
    iteration counts as the cycle launched kernel by kernel; so does one
    preconditioner application."""
    d, spaces, mats, rhs = levels(ctx, 4, 3)
    fs, A, b = spaces[-1], mats[-1], rhs[-1]
    out = {}
    for graph in (0, 1):
        x0 = torch.zeros_like(b)
        x0[fs.active_bound:] = b[fs.active_bound:]
        prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
                               residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=300,
                               multigrid_cycle_type="V", multigrid_smoother="jacobi",
                               multigrid_smoother_coarse="direct_solve", multigrid_n_pre_smooth=2,
                               multigrid_n_post_smooth=2, jacobi_damp=0.6, precision=precision,
                               graph_cycle=graph)
        sol = pm.Solver(ctx, prm)
        sol.update_matrix(mats, spaces)
        runs = []
        for rep in range(3):
            x = x0.clone()
            r = sol.solve(b, x)
            z = torch.zeros_like(b)
            sol.apply_preconditioner(b.clone(), z)
            y = torch.zeros_like(b)
            A.apply(x, y)
            runs.append((x.cpu().numpy(), r.iterations, z.cpu().numpy()))
        out[graph] = runs
    for k in range(3):
        assert np.array_equal(out[0][k][0], out[1][k][0]) and out[0][k][1] == out[1][k][1], k
        assert np.array_equal(out[0][k][2], out[1][k][2]), k
        assert np.array_equal(out[1][0][0], out[1][k][0]), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_value_update_equals_a_fresh_set_up(ctx, precision):
    """Solver::update_matrix with the same matrices and spaces as before only
    recomputes what depends on the values (diagonals, bf16 copies, FP32
    copies, coarse inverse) in place: after the matrices are reassembled with
    other coefficients, the refreshed solver and a freshly set-up one give the
    same bits and iteration counts."""
    d, spaces, mats, rhs = levels(ctx, 4, 3)
    prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
                           residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=300,
                           multigrid_cycle_type="V", multigrid_smoother="jacobi",
                           multigrid_smoother_coarse="direct_solve", multigrid_n_pre_smooth=2,
                           multigrid_n_post_smooth=2, jacobi_damp=0.6, precision=precision)
    sol = pm.Solver(ctx, prm)
    sol.update_matrix(mats, spaces)
    fs = spaces[-1]
    x = torch.zeros_like(rhs[-1])
    x[fs.active_bound:] = rhs[-1][fs.active_bound:]
    sol.solve(rhs[-1], x.clone())
    kw2 = dict(KW, eps=5e-3, b=(0.3, -0.8, 0.5))
    for s, A, f in zip(spaces, mats, rhs):
        A.reset()
        f.zero_()
        pm.assemble_cd(ctx, s, pm.cd_params(**kw2), A, f)
    b = rhs[-1]
    x0 = torch.zeros_like(b)
    x0[fs.active_bound:] = b[fs.active_bound:]
    sol.update_matrix(mats, spaces)          # same hierarchy: values only
    x1 = x0.clone()
    r1 = sol.solve(b, x1)
    fresh = pm.Solver(ctx, prm)
    fresh.update_matrix(mats, spaces)
    x2 = x0.clone()
    r2 = fresh.solve(b, x2)
    assert r1.converged and r1.iterations == r2.iterations
    assert torch.equal(x1, x2)


def test_direct_coarse_solve_with_pivoting_matches_oracle(ctx):
    """The blocked Gauss-Jordan inverse of the coarsest matrix (729 DOFs:
    eleven 64-column blocks and a partial one) in a 2-level cycle against the
    oracle's cycle with its LU solve, at 1e-10."""
    nref = 3
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    spaces = [pm.TFESpace(d, l, 2, ctx) for l in (nref - 1, nref)]
    mats = []
    for s in spaces:
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.assemble_cd(ctx, s, pm.cd_params(**KW), A, f)
        mats.append(A)
    prm = pm.solver_params(iterative_solver_type="fgmres", preconditioner="multigrid", multigrid_cycle_type="V",
                           multigrid_smoother="jacobi", multigrid_smoother_coarse="direct_solve",
                           multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2, jacobi_damp=0.6)
    sol = pm.Solver(ctx, prm)
    sol.update_matrix(mats, spaces)
    rng = np.random.default_rng(2)
    v = torch.as_tensor(rng.standard_normal(spaces[-1].n_dof)).cuda()
    z = torch.zeros_like(v)
    sol.apply_preconditioner(v, z)
    osp = [O.build_space(2, nref, l, 2)[0] for l in (nref - 1, nref)]
    op = O.solver_params(preconditioner=2, mg_cycle=0, mg_smoother_coarse=1, mg_n_pre_smooth=2, mg_n_post_smooth=2,
                         jacobi_damp=0.6, mg_correction_damp=1.0)
    zo = O.mg_apply(op, osp, [A.get_entries() for A in mats], v.cpu().numpy())
    assert spaces[0].n_dof == 729
    assert np.max(np.abs(z.cpu().numpy() - zo)) <= 1e-10 * np.max(np.abs(zo))
