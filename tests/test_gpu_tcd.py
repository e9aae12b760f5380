"""Time-dependent convection-diffusion on the HIP path (through the C ABI):
TCD assembly against the oracle, the reference's tcd2d_test golden errors
(Crank-Nicolson, Galerkin and SUPG) and the exactness of the linear
space-time example for every scheme."""
import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEO = {"UnitSquare": 0, "TwoTriangles": 1}


@pytest.mark.parametrize("geo,order", [("UnitSquare", 1), ("TwoTriangles", 1), ("UnitSquare", 2)])
@pytest.mark.parametrize("disc", ["galerkin", "supg"])
def test_tcd_assembly_matches_oracle(ctx, geo, order, disc):
    print("case", geo, order, disc, flush=True)
    nref = 3
    kw = dict(eps=0.01, b=(1.0, -0.5), c=0.7)
    d = pm.TDomain(geo, nref)
    s = pm.TFESpace(d, nref, order, ctx)
    A, M = pm.FEMatrix(ctx, s), pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    p = pm.cd_params("tcd_exp", disc=disc, time=0.3, theta1=0.5, tau=0.05, **kw)
    pm.assemble_tcd(ctx, s, p, A, M, f)
    sp, info = O.build_space(GEO[geo], nref, nref, order)
    op = O.cd_params(example=7, disc=1 if disc == "supg" else 0, eps=kw["eps"], b=kw["b"] + (0.0,),
                     c=kw["c"], time=0.3, theta1=0.5, tau=0.05)
    a, m, r = O.assemble_tcd(sp, op)
    for got, ref in ((A.get_entries(), a), (M.get_entries(), m)):
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.max(np.abs(f.cpu().numpy() - r)) <= 1e-12 * np.max(np.abs(r))
    # stiffness and rhs only (the TCDStiffRhs assembling of a Galerkin time step)
    A.reset()
    f.zero_()
    pm.assemble_tcd(ctx, s, p, A, None, f)
    assert np.max(np.abs(A.get_entries() - a)) <= 1e-12 * np.max(np.abs(a))


def fgmres_jacobi_tight():
    return pm.solver_params(iterative_solver_type="fgmres", preconditioner="jacobi", gmres_restart=50,
                            max_n_iterations=3000, residual_tolerance=1e-13, residual_reduction=0.0)


def integrate(ctx, geo, nref, order, example, disc, scheme, tau, n_steps, **kw):
    d = pm.TDomain(geo, nref)
    tss = pm.TimeDiscretization(ctx, scheme, tau, 0.0, tau * n_steps)
    tcd = pm.TimeConvectionDiffusion(ctx, d, order, example, tss, fgmres_jacobi_tight(), disc=disc, **kw)
    tss.current_step_ = 0
    tcd.assemble_initial_time()
    errs = [tcd.output()[:2]]
    while not tss.reached_final_time_step():
        tss.current_step_ += 1
        tss.current_time_ += tss.tau
        tcd.assemble()
        rep = tcd.solve()
        assert rep.converged, (tss.current_step_, rep.iterations)
        errs.append(tcd.output()[:2])
    return np.array(errs), tcd


@pytest.mark.parametrize("disc", ["galerkin", "supg"])
def test_tcd2d_golden_errors_on_gpu(ctx, golden, disc):
    """test/tcd2d_test.c++ on the GPU: Crank-Nicolson, Q1 on 32x32 quads."""
    g = golden["tcd2d_test"]
    errs, _ = integrate(ctx, "UnitSquare", 5, 1, "tcd_exp", disc, "crank_nicolson", g["time_step_length"],
                        g["n_steps"], eps=1.0, b=(1.0, -1.0, 0.0), c=1.0)
    assert len(errs) == g["n_steps"] + 1
    for k, m in enumerate(g["steps"]):
        assert abs(errs[m][0] - g[disc]["L2"][k]) < g["tolerance"], (m, errs[m][0], g[disc]["L2"][k])
        assert abs(errs[m][1] - g[disc]["H1"][k]) < g["tolerance"], (m, errs[m][1], g[disc]["H1"][k])


@pytest.mark.parametrize("scheme", ["backward_euler", "crank_nicolson", "bdf_two"])
def test_tcd_linear_space_time_exact_on_gpu(ctx, scheme):
    errs, _ = integrate(ctx, "TwoTriangles", 3, 1, "tcd_linear", "galerkin", scheme, 0.1, 5,
                        eps=1.0, b=(1.0, 2.0, 0.0), c=1.0)
    assert np.max(errs) < 1e-10, errs


def test_tcd_backward_euler_matches_oracle(ctx):
    """Solution after 5 backward-Euler steps (P1, SUPG) against the oracle's
    restatement of TimeConvectionDiffusion."""
    kw = dict(eps=1e-3, b=(np.cos(0.7), np.sin(0.7), 0.0), c=0.0)
    errs, tcd = integrate(ctx, "TwoTriangles", 4, 1, "tcd_exp", "supg", "backward_euler", 0.01, 5, **kw)
    sp, info = O.build_space(1, 4, 4, 1)
    oerrs, u = O.tcd_time_integration(sp, 7, 1, O.BACKWARD_EULER, 0.01, 5, eps=kw["eps"], b=kw["b"][:2],
                                      c=kw["c"])
    ug = tcd.solution.cpu().numpy()
    assert np.max(np.abs(ug - u)) <= 1e-10 * np.max(np.abs(u))
    assert np.allclose(errs, oerrs, rtol=1e-9, atol=1e-14)
