"""The C ABI from a C++ program: tests/host/pmx_caller.cpp, compiled with g++
against include/pmx.h alone and linked with libpmx.so (the shape of the
reference-side adapter), runs a 3-level CD3D problem whose coefficients and
boundary values it computes itself (dev_coeff_table, host_dirichlet), solves
it and checks the error path; its results must agree with the same problem
through the Python mirror and the built-in example."""
import json
import os
import subprocess

import numpy as np
import pytest
import torch

from mymoonmd_amd import build as B
from mymoonmd_amd import parmoon as pm

pytestmark = pytest.mark.gpu

B3 = (1 / 14 ** 0.5, 2 / 14 ** 0.5, 3 / 14 ** 0.5)


def test_cpp_caller_matches_python_route(ctx):
    if not os.path.exists(B.CALLER):
        pytest.fail("tests/host/pmx_caller is not built: run python -m mymoonmd_amd.build")
    nref = 3
    print("pmx_caller", nref, flush=True)
    r = subprocess.run([B.CALLER, str(nref)], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["failures"] == 0 and out["matrix_equal"], out
    assert out["rhs_rel_diff"] <= 1e-13 and out["fp64_rel_residual"] <= 1.01e-8
    assert out["gemm_route_rel_diff"] <= 1e-13
    assert out["error_code"] == 1 and "NULL" in out["error_message"]
    # the Python mirror on the built-in example
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    spaces = [pm.TFESpace(d, l, 2, ctx) for l in range(nref - 2, nref + 1)]
    prm = pm.cd_params("exp_sin_3d", disc="supg", eps=1e-3, b=B3, hexa_quad_rule=2)
    mats, rhss = [], []
    for s in spaces:
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.assemble_cd(ctx, s, prm, A, f)
        mats.append(A)
        rhss.append(f)
    sol = pm.Solver(ctx, pm.solver_params(
        iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30, max_n_iterations=200,
        residual_tolerance=0.0, residual_reduction=1e-8, multigrid_cycle_type="V", multigrid_smoother="jacobi",
        multigrid_smoother_coarse="direct_solve", multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2,
        jacobi_damp=0.7, precision="mixed"))
    sol.update_matrix(mats, spaces)
    b = rhss[-1]
    ab = spaces[-1].active_bound
    x = torch.zeros_like(b)
    x[ab:] = b[ab:]
    rep = sol.solve(b, x)
    err = pm.compute_errors(ctx, spaces[-1], prm, x)
    assert out["n_dof"] == spaces[-1].n_dof
    assert abs(out["iterations"] - rep.iterations) <= 1, (out["iterations"], rep.iterations)
    assert abs(out["l2"] - err[0]) <= 1e-6 * err[0] and abs(out["h1"] - err[1]) <= 1e-6 * err[1], (out, err)
