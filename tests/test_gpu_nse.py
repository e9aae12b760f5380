"""Oseen / Navier-Stokes (Q2/Q1, NSTYPE 1) on the HIP path against the oracle:
numbering, patterns, assembled blocks, block operator, combined matrix,
interpolation, errors and the pressure projection."""
import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture(scope="module", params=[1, 2])
def pair(ctx, request):
    nref = request.param
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    g = pm.NSESystem(ctx, d, nref)
    o = O.NseOracle(nref, nref)
    return nref, d, g, o


def test_nse_numbering_and_patterns(pair):
    nref, d, g, o = pair
    gv, gp = g.global_numbers()
    assert np.array_equal(gv, o.vs.gnum) and np.array_equal(gp, o.ps.gnum)
    rp, col = g.structure("A")
    assert np.array_equal(rp, o.vs.rp) and np.array_equal(col, o.vs.col)
    rp, col = g.structure("B")
    assert np.array_equal(rp, o.brp) and np.array_equal(col, o.bcol)


@pytest.mark.parametrize("nu,conv", [(1.0, False), (1e-2, True)])
def test_nse_assembly_apply_combined_match_oracle(ctx, pair, nu, conv):
    nref, d, g, o = pair
    print("case", nref, nu, conv, flush=True)
    rng = np.random.default_rng(7)
    w = rng.standard_normal(3 * o.n_u) if conv else None
    rhs = g.zeros()
    g.assemble("u2_p1", nu, dev(w) if conv else None, rhs)
    A, B, r = o.assemble(nu, w)
    ga = g.values("A")
    assert np.max(np.abs(ga - A)) <= 1e-12 * np.max(np.abs(A))
    for k in range(3):
        gb = g.values("B%d" % (k + 1))
        assert np.max(np.abs(gb - B[k])) <= 1e-12 * np.max(np.abs(B[k]))
    assert np.max(np.abs(rhs.cpu().numpy() - r)) <= 1e-12 * np.max(np.abs(r))
    x = rng.standard_normal(o.n)
    y = g.zeros()
    g.apply(dev(x), y)
    yo = o.apply(A, B, x)
    assert np.max(np.abs(y.cpu().numpy() - yo)) <= 1e-12 * np.max(np.abs(yo))
    K = o.combined(A, B)
    rp, col = g.structure("combined")
    import scipy.sparse as sp
    G = sp.csr_matrix((g.values("combined"), col, rp), shape=(o.n, o.n))
    assert abs(G - K).max() <= 1e-12 * abs(K).max()


def test_nse_interpolant_residual_errors_and_projection(ctx, pair):
    """The interpolant of test_u_2_p_1 solves the Oseen system with w = u; its
    errors vanish; the pressure projection removes a constant shift."""
    nref, d, g, o = pair
    x = g.zeros()
    g.interpolate(x)
    assert np.max(np.abs(x.cpu().numpy() - O.nse_interpolate(o))) < 1e-14
    rhs = g.zeros()
    g.assemble("u2_p1", 1e-2, x[:3 * g.n_u].clone(), rhs)
    r = rhs.clone()
    g.apply_scaled_add(x, r, -1.0)
    assert torch.max(torch.abs(r)).item() < 1e-13
    e = g.errors(x)
    assert e[0] < 1e-13 and e[1] < 1e-12 and e[3] < 1e-13 and e[4] < 1e-12, e
    x[3 * g.n_u:] += 0.37
    assert g.errors(x)[3] > 0.3
    g.project_pressure(x)
    assert g.errors(x)[3] < 1e-13


MG_KW = dict(iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
             multigrid_cycle_type="V", multigrid_smoother="cell_vanka_jacobi",
             multigrid_smoother_coarse="direct_solve", multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2,
             multigrid_vanka_damp_factor=0.5, multigrid_correction_damp_factor=0.8)


def two_levels(ctx, nref, nu=1e-2):
    """GPU and oracle NSE systems of levels nref-1, nref with w = interpolant of u."""
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    gl, ol, rhs = [], [], None
    for lev in (nref - 1, nref):
        g = pm.NSESystem(ctx, d, lev)
        x = g.zeros()
        g.interpolate(x)
        r = g.zeros()
        g.assemble("u2_p1", nu, x[:3 * g.n_u].clone(), r)
        o = O.NseOracle(nref, lev)
        xo = O.nse_interpolate(o)
        A, B, ro = o.assemble(nu, xo[:3 * o.n_u])
        gl.append(g)
        ol.append(O.NseLevelOracle(o, A, B))
        rhs = (r, ro, x)
    return d, gl, ol, rhs


@pytest.mark.parametrize("coarse", ["direct_solve", "cell_vanka_jacobi"])
def test_nse_multigrid_vanka_cycle_matches_oracle(ctx, coarse):
    d, gl, ol, (r, ro, xe) = two_levels(ctx, 2)
    kw = dict(MG_KW, multigrid_smoother_coarse=coarse, multigrid_coarse_max_n_iterations=3,
              multigrid_coarse_residual=1e-30)
    sol = pm.Solver(ctx, pm.solver_params(**kw))
    sol.update_saddle(gl)
    v = np.random.default_rng(3).standard_normal(gl[-1].n)
    z = gl[-1].zeros()
    sol.apply_preconditioner(dev(v), z)
    zo = O.nse_mg_apply(ol, v, damp_corr=0.8, vanka_damp=0.5, n_pre=2, n_post=2, coarse=coarse,
                        coarse_maxit=3, coarse_res=1e-30)
    assert np.max(np.abs(z.cpu().numpy() - zo)) <= 1e-10 * np.max(np.abs(zo))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nse_fgmres_multigrid_vanka_solves_example(ctx, precision):
    """test_u_2_p_1 (NSE3D example -3, Q2/Q1): the Oseen system with w = u is
    solved exactly (errors of the interpolant ~ 0), in as many FGMRES
    iterations as the oracle (FP64 Vanka factors) or close to it (FP32)."""
    d, gl, ol, (r, ro, xe) = two_levels(ctx, 2)
    g = gl[-1]
    sol = pm.Solver(ctx, pm.solver_params(residual_tolerance=1e-11, residual_reduction=0.0, max_n_iterations=200,
                                          precision=precision, **MG_KW))
    sol.update_saddle(gl)
    x = g.zeros()
    nu, na = g.n_u, g.n_u_active
    for k in range(3):
        x[k * nu + na:(k + 1) * nu] = r[k * nu + na:(k + 1) * nu]
    rep = sol.solve(r, x)
    assert rep.converged
    g.project_pressure(x)
    e = g.errors(x)
    assert e[0] < 1e-9 and e[1] < 1e-8 and e[3] < 1e-8, e
    F = ol[-1]
    x0 = np.zeros(F.ns.n)
    for k in range(3):
        x0[k * nu + na:(k + 1) * nu] = ro[k * nu + na:(k + 1) * nu]
    xo, it = O.fgmres(lambda z: F.ns.apply(F.A, F.B, z),
                      lambda v: O.nse_mg_apply(ol, v, damp_corr=0.8, vanka_damp=0.5, n_pre=2, n_post=2), ro, x0,
                      restart=30, max_it=200, tol=1e-11)
    if precision == "fp64":
        assert abs(rep.iterations - it) <= 1, (rep.iterations, it)
    else:
        assert rep.iterations <= it + 5, (rep.iterations, it)


def gpu_levels(ctx, nref, nlev, nu=1e-2):
    """GPU NSE systems of the nlev finest levels up to nref, w = interpolant of u."""
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    gl, r = [], None
    for lev in range(nref - nlev + 1, nref + 1):
        g = pm.NSESystem(ctx, d, lev)
        x = g.zeros()
        g.interpolate(x)
        r = g.zeros()
        g.assemble("u2_p1", nu, x[:3 * g.n_u].clone(), r)
        gl.append(g)
    return d, gl, r


def test_nse_three_level_cycle_matches_oracle(ctx):
    """Levels 1, 2, 3 (8, 64, 512 hexahedra): the recursion over more than one
    coarse grid against the oracle's cycle (level 0, one hexahedron, has a
    singular saddle-point matrix)."""
    d, gl, r = gpu_levels(ctx, 3, 3)
    ol = []
    for lev in range(1, 4):
        o = O.NseOracle(3, lev)
        A, B, _ = o.assemble(1e-2, O.nse_interpolate(o)[:3 * o.n_u])
        ol.append(O.NseLevelOracle(o, A, B))
    sol = pm.Solver(ctx, pm.solver_params(**dict(MG_KW, multigrid_vanka_damp_factor=0.3)))
    sol.update_saddle(gl)
    v = np.random.default_rng(11).standard_normal(gl[-1].n)
    z = gl[-1].zeros()
    sol.apply_preconditioner(dev(v), z)
    zo = O.nse_mg_apply(ol, v, damp_corr=0.8, vanka_damp=0.3, n_pre=2, n_post=2)
    assert np.max(np.abs(z.cpu().numpy() - zo)) <= 1e-10 * np.max(np.abs(zo))


@pytest.mark.parametrize("nref,nlev", [(3, 3), (4, 4)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nse_multilevel_vanka_converges(ctx, nref, nlev, precision):
    """8^3 and 16^3 hexahedra on 3 and 4 levels: with the Vanka damping 0.3
    (cell-Vanka with continuous Q1 pressure needs damping below 0.5 on the
    finer grids) FGMRES + multigrid reaches 1e-8 in a bounded number of
    iterations and the solution is the example's interpolant."""
    print("case", nref, nlev, precision, flush=True)
    d, gl, r = gpu_levels(ctx, nref, nlev)
    g = gl[-1]
    kw = dict(MG_KW, multigrid_vanka_damp_factor=0.3)
    sol = pm.Solver(ctx, pm.solver_params(residual_tolerance=0.0, residual_reduction=1e-10, max_n_iterations=200,
                                          precision=precision, **kw))
    sol.update_saddle(gl)
    x = g.zeros()
    nu, na = g.n_u, g.n_u_active
    for k in range(3):
        x[k * nu + na:(k + 1) * nu] = r[k * nu + na:(k + 1) * nu]
    rep = sol.solve(r, x)
    print("iterations", rep.iterations, flush=True)
    assert rep.converged and rep.iterations <= 80, rep.iterations
    g.project_pressure(x)
    e = g.errors(x)
    assert e[0] < 1e-8 and e[3] < 1e-7, e


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nse_vanka_local_inverse_matches_lu(ctx, precision):
    """The Jacobi cell-Vanka smoother with the explicit local inverses (built
    in place from the LU factors) against the one substituting with the LU
    factors: one cycle agrees to rounding (1e-11 FP64, 1e-5 with FP32 local
    matrices) and FGMRES takes the same number of iterations to the same
    solution."""
    d, gl, r = gpu_levels(ctx, 3, 3)
    g = gl[-1]
    kw = dict(MG_KW, multigrid_vanka_damp_factor=0.3)
    v = dev(np.random.default_rng(5).standard_normal(g.n))
    out = {}
    for inv in (0, 1):
        sol = pm.Solver(ctx, pm.solver_params(residual_tolerance=0.0, residual_reduction=1e-10, max_n_iterations=200,
                                              precision=precision, vanka_inverse=inv, **kw))
        sol.update_saddle(gl)
        z = g.zeros()
        sol.apply_preconditioner(v.clone(), z)
        x = g.zeros()
        nu, na = g.n_u, g.n_u_active
        for k in range(3):
            x[k * nu + na:(k + 1) * nu] = r[k * nu + na:(k + 1) * nu]
        rep = sol.solve(r, x)
        out[inv] = (z.cpu().numpy(), x.cpu().numpy(), rep.iterations)
    tol = 1e-11 if precision == "fp64" else 1e-5
    a, b = out[0], out[1]
    assert np.max(np.abs(a[0] - b[0])) <= tol * np.max(np.abs(a[0]))
    assert abs(a[2] - b[2]) <= 1 and np.max(np.abs(a[1] - b[1])) <= 1e-7 * np.max(np.abs(a[1]))


def test_nse_restrict_velocity_matches_oracle(ctx):
    """GridTransfer::RestrictFunction of the three velocity components (4^3 ->
    2^3 hexahedra) against the oracle, and the restriction of the example's
    quadratic velocity is its coarse interpolant."""
    d = pm.TDomain("Default_UnitCube_Hexa", 2)
    gc, gf = pm.NSESystem(ctx, d, 1), pm.NSESystem(ctx, d, 2)
    oc, of = O.NseOracle(2, 1), O.NseOracle(2, 2)
    u = np.random.default_rng(5).standard_normal(3 * gf.n_u)
    out = torch.zeros(3 * gc.n_u, dtype=torch.float64, device="cuda")
    gc.restrict_velocity(gf, dev(u), out)
    got = out.cpu().numpy()
    for k in range(3):
        ref = O.restrict_function(oc.vs, of.vs, u[k * gf.n_u:(k + 1) * gf.n_u])
        assert np.max(np.abs(got[k * gc.n_u:(k + 1) * gc.n_u] - ref)) <= 1e-12 * np.max(np.abs(ref))
    xf, xc = gf.zeros(), gc.zeros()
    gf.interpolate(xf)
    gc.interpolate(xc)
    gc.restrict_velocity(gf, xf, out)
    assert torch.max(torch.abs(out - xc[:3 * gc.n_u])).item() < 1e-13


def test_nse_picard_matches_oracle(ctx):
    """The nonlinear loop of NavierStokes<3> (Picard: convection with the last
    velocity, restricted to the coarse level) on 4^3 hexahedra with two
    multigrid levels: the residual history follows the oracle's and the loop
    ends at the example's interpolant (the exact discrete solution)."""
    d = pm.TDomain("Default_UnitCube_Hexa", 2)
    prm = pm.solver_params(residual_tolerance=1e-11, residual_reduction=0.0, max_n_iterations=200, **MG_KW)
    ns = pm.NavierStokes(ctx, d, [1, 2], prm, nu=0.1, nonlinloop_epsilon=1e-10, nonlinloop_maxit=20)
    ns.assemble_linear_terms()
    ns.stop_it(0)
    k = 0
    while True:
        k += 1
        ns.solve()
        ns.assemble_nonlinear_term()
        if ns.stop_it(k):
            break
    hist = [r[2] for r in ns.residuals]
    kw = dict(damp_corr=0.8, vanka_damp=0.5, n_pre=2, n_post=2)
    xo, ho, its = O.nse_picard(2, [1, 2], 0.1, kw, eps=1e-10, max_it=20)
    print("gpu", hist, ns.linear_iterations, "oracle", ho, its, flush=True)
    assert len(hist) == len(ho)
    for a, b in zip(hist, ho):
        if b > 1e-8:
            assert abs(a - b) <= 1e-6 * b, (a, b)
    f = ns.systems[-1]
    e = f.errors(ns.solution[-1])
    assert e[0] < 1e-9 and e[3] < 1e-8, e


@pytest.mark.parametrize("coarse", ["direct_solve", "cell_vanka"])
def test_nse_multiplicative_vanka_cycle_matches_oracle(ctx, coarse):
    """VankaSmoother CELL (multiplicative) in colour-major cell order: one
    V(2,2) cycle on 2 levels against the oracle's sequential sweep in the same
    order (the GPU runs the 27 colours one after another, the cells of a
    colour in parallel)."""
    d, gl, ol, (r, ro, xe) = two_levels(ctx, 2)
    kw = dict(MG_KW, multigrid_smoother="cell_vanka", multigrid_smoother_coarse=coarse,
              multigrid_coarse_max_n_iterations=3, multigrid_coarse_residual=1e-30)
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-5)):
        sol = pm.Solver(ctx, pm.solver_params(precision=precision, **kw))
        sol.update_saddle(gl)
        v = np.random.default_rng(3).standard_normal(gl[-1].n)
        z = gl[-1].zeros()
        sol.apply_preconditioner(dev(v), z)
        zo = O.nse_mg_apply(ol, v, damp_corr=0.8, vanka_damp=0.5, n_pre=2, n_post=2, coarse=coarse,
                            coarse_maxit=3, coarse_res=1e-30, smoother="cell_vanka")
        assert np.max(np.abs(z.cpu().numpy() - zo)) <= tol * np.max(np.abs(zo)), precision


def test_nse_multiplicative_vanka_solves_faster_than_jacobi_variant(ctx):
    """On 8^3 cells with 3 levels both Vanka variants solve the example; the
    multiplicative one needs fewer FGMRES iterations (and tolerates the
    damping 0.5 at which the averaged variant needs 0.3)."""
    d, gl, r = gpu_levels(ctx, 3, 3)
    g = gl[-1]
    its = {}
    for smoother, damp in (("cell_vanka", 0.5), ("cell_vanka_jacobi", 0.3)):
        kw = dict(MG_KW, multigrid_smoother=smoother, multigrid_vanka_damp_factor=damp)
        sol = pm.Solver(ctx, pm.solver_params(residual_tolerance=0.0, residual_reduction=1e-10,
                                              max_n_iterations=200, **kw))
        sol.update_saddle(gl)
        x = g.zeros()
        nu, na = g.n_u, g.n_u_active
        for k in range(3):
            x[k * nu + na:(k + 1) * nu] = r[k * nu + na:(k + 1) * nu]
        rep = sol.solve(r, x)
        assert rep.converged, smoother
        g.project_pressure(x)
        e = g.errors(x)
        assert e[0] < 1e-8 and e[3] < 1e-7, (smoother, e)
        its[smoother] = rep.iterations
    print("iterations", its, flush=True)
    assert its["cell_vanka"] < its["cell_vanka_jacobi"], its


def test_nse_vanka_bf16_local_inverses(ctx):
    """Mixed precision with vanka_inverse 2 (bf16 local inverses and bf16
    values in the smoother's defect) against vanka_inverse 1 (FP32): one cycle
    differs only by bf16 rounding (relative 2e-2 at most), and FGMRES reaches
    the same solution (1e-7, pressures of mean zero) in at most one more
    iteration."""
    d, gl, r = gpu_levels(ctx, 3, 3)
    g = gl[-1]
    kw = dict(MG_KW, multigrid_vanka_damp_factor=0.3)
    v = dev(np.random.default_rng(7).standard_normal(g.n))
    out = {}
    for inv in (1, 2):
        print("case vanka_inverse", inv, flush=True)
        sol = pm.Solver(ctx, pm.solver_params(residual_tolerance=0.0, residual_reduction=1e-10, max_n_iterations=200,
                                              precision="mixed", vanka_inverse=inv, **kw))
        sol.update_saddle(gl)
        z = g.zeros()
        sol.apply_preconditioner(v.clone(), z)
        x = g.zeros()
        nu, na = g.n_u, g.n_u_active
        for k in range(3):
            x[k * nu + na:(k + 1) * nu] = r[k * nu + na:(k + 1) * nu]
        rep = sol.solve(r, x)
        assert rep.converged
        g.project_pressure(x)   # the pressure is fixed up to a constant
        out[inv] = (z.cpu().numpy(), x.cpu().numpy(), rep.iterations)
    a, b = out[1], out[2]
    dz = np.max(np.abs(a[0] - b[0])) / np.max(np.abs(a[0]))
    assert 0 < dz <= 2e-2, dz
    assert b[2] <= a[2] + 1, (a[2], b[2])
    assert np.max(np.abs(a[1] - b[1])) <= 1e-7 * np.max(np.abs(a[1]))
