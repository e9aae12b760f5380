"""Benchmark of the CD3D solve path on one MI355X.

Workload (the reference metric's configuration): unit cube, 64^3 Q2 hexahedra
(2,146,689 DOFs, 130,422,149 nnz), eps = 1e-3, b = (1,2,3)/sqrt(14), c = 0,
SUPG, Gauss 3x3x3, Dirichlet data and right-hand side of
u = exp(-x) sin(pi y) sin(pi z); FGMRES(30) preconditioned by a 4-level
geometric multigrid V(2,2) with damped Jacobi, stopping at a relative residual
of 1e-8 (residual_reduction 1e-8, residual_tolerance 0); the coarsest level
(8^3 cells, 4,913 DOFs) is solved exactly (the reference's default
multigrid_smoother_coarse = direct_solve).

Jacobi damping: with omega = 0.8 the smoother is not contractive on this SUPG
operator (the largest eigenvalue of D^-1 A is about 3.04, so |1 - 0.8*3.04| =
1.43) and FGMRES stagnates near 1e-5 relative residual (6000 iterations do not
reach 1e-8).  The timed solves therefore use omega = 0.7, the largest damping
in steps of 0.1 that reaches 1e-8 (35 iterations); the omega = 0.8 solve is run
once with the iteration cap and its residual is reported beside the timings.

One step = one complete solve from the Dirichlet-initialised guess with all
inputs resident in HBM.  `value` is the time to solution of the mixed path
(FP32 V-cycle on bf16 row-scaled matrices inside FP64 FGMRES); the plain FP64
path is timed in the same run.  Every result is checked in FP64 on the GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B = (1 / math.sqrt(14), 2 / math.sqrt(14), 3 / math.sqrt(14))
EPS = 1e-3
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--refine", type=int, default=6, help="cube refinements (6 -> 64^3)")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--jacobi-damp", type=float, default=0.7)
    ap.add_argument("--damp08-iterations", type=int, default=300,
                    help="iteration cap of the omega = 0.8 check (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched TCD2D sweep")
    ap.add_argument("--no-oseen", action="store_true", help="skip the Oseen solves (Q2/Q1 and Q2/P1disc)")
    ap.add_argument("--profile-kernel", default=None,
                    choices=["apply_fp64", "jacobi_fp64", "apply_bf16", "jacobi_bf16", "solver_jacobi_bf16",
                             "solver_jacobi_fp64", "assembly", "vanka", "vanka_p1disc", "batch", "batch_gmres"],
                    help="only run this workload a few times (an operator kernel 5 times, the multigrid's "
                         "finest sweep, the 64^3 assembly, the Oseen Vanka solve, the batched sweep) and "
                         "copy 1 GiB 3 times (the calibration of HBM byte counters), then exit")
    ap.add_argument("--cpu-refine", type=int, default=4, help="refinements of the CPU sample")
    ap.add_argument("--renumber-levels", type=int, default=None,
                    help="solver parameter renumber_levels (default: the solver's default)")
    ap.add_argument("--inner-reduction", type=float, default=None,
                    help="solver parameter refinement_inner_reduction of the mixed path (default: the solver's)")
    return ap.parse_args()


def solver_params(pm, precision, omega, max_it=500, renumber=None, inner=None, reduction=1e-8):
    extra = {} if renumber is None else {"renumber_levels": renumber}
    if inner is not None:
        extra["refinement_inner_reduction"] = inner
    return pm.solver_params(**extra,
        iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30,
        residual_tolerance=0.0, residual_reduction=reduction, max_n_iterations=max_it,
        multigrid_cycle_type="V", multigrid_smoother="jacobi", multigrid_smoother_coarse="direct_solve",
        multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2, jacobi_damp=omega,
        multigrid_correction_damp_factor=1.0, multigrid_coarse_residual=0.1,
        multigrid_coarse_max_n_iterations=10, precision=precision)


def bf16_row_scaled_error(vals, row_ptr):
    """max |bf16(a / s) s - a| / s over the entries of each row, s = max |a| of
    the row: the error of the multigrid's bf16 row-scaled copy (rounding to
    nearest even, as the device fill does)"""
    import numpy as np
    v = np.asarray(vals, np.float64)
    rp = np.asarray(row_ptr, np.int64)
    lens = np.diff(rp)
    nz = lens > 0
    scale = np.zeros(len(lens))
    scale[nz] = np.maximum.reduceat(np.abs(v), rp[:-1][nz])
    srow = np.repeat(scale, lens)
    srow[srow == 0] = 1.0
    f = (v / srow).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    back = u.astype(np.uint32).view(np.float32).astype(np.float64)
    return float(np.max(np.abs(back - v / srow)))


def folded_entries(A, active):
    """(active rows, entries of active rows in active columns) of a matrix"""
    import numpy as np
    rp, col = A.get_structure()
    end = int(rp[active])
    return active, end - int(np.count_nonzero(col[:end] >= active))


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at 64^3 from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by the copy-kernel calibration it
    states), or None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "sell_spmv_cd3d_q2_64_pmc_traffic.txt")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch %s " % kernel):
                    return int(line.split()[2])
    except OSError:
        pass
    return None


def cpu_baseline(nref, nlev, omega, max_it=60):
    """The oracle (sequential C restatement of the reference CPU path) on a
    bounded sample of the same problem, one host core."""
    import numpy as np
    from oracle import oracle as O
    spaces = []
    for l in range(nref - nlev + 1, nref + 1):
        sp, _ = O.build_space(2, nref, l, 2)
        spaces.append(sp)
    p = O.cd_params(example=5, disc=1, eps=EPS, b=B, hexa_rule=2)
    t0 = time.perf_counter()
    vals, rhss = [], []
    for sp in spaces:
        v, r = O.assemble(sp, p)
        vals.append(v)
        rhss.append(r)
    t_asm_all = time.perf_counter() - t0
    fs = spaces[-1]
    t0 = time.perf_counter()
    vf, rf = O.assemble(fs, p)
    y = np.zeros(fs.n_dof)
    O.apply_scaled_add(fs.n_dof, fs.active, fs.rp, fs.col, vf, rf, y, 1.0)
    t_asm_spmv = time.perf_counter() - t0
    x0 = np.zeros(fs.n_dof)
    x0[fs.active:] = rf[fs.active:]
    op = O.solver_params(iterative_solver_type=0, preconditioner=2, gmres_restart=30,
                         residual_tolerance=0.0, residual_reduction=1e-8, max_n_iterations=max_it,
                         mg_cycle=0, mg_smoother_coarse=1, mg_n_pre_smooth=2, mg_n_post_smooth=2,
                         jacobi_damp=omega, mg_correction_damp=1.0, mg_coarse_residual=0.1,
                         mg_coarse_max_n_iterations=10)
    t0 = time.perf_counter()
    x, it, hist, st = O.solve(op, spaces, vals, rf, x0)
    t_solve = time.perf_counter() - t0
    return dict(n_dof=fs.n_dof, t_solve=t_solve, iterations=it, t_asm_spmv=t_asm_spmv,
                t_asm_all=t_asm_all, mdofs=fs.n_dof / t_asm_spmv / 1e6)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_full(pm, nref, spaces, mats, b, x0, omega, iterations=(1, 2)):
    """The oracle (sequential C restatement of the reference CPU path, one
    core) on the full 64^3 problem: FGMRES(30) + 4-level MG V(2,2) on the
    GPU-assembled FP64 matrices (equal to the oracle's assembly at 1e-12) with
    the numbering and pattern of the product (equal to the oracle's, bit for
    bit), coarse level smoothed by 10 Jacobi steps (the oracle's dense LU of
    the 4,913 coarse DOFs alone takes tens of seconds; the reference uses a
    sparse LU there).  Two runs of 1 and 2 iterations give the time per
    full-size iteration without the set-up."""
    import numpy as np
    from oracle import oracle as O
    shape, lv = O.mesh(2, nref)
    osp = []
    for s, M in zip(spaces, mats):
        rp, col = M.get_structure()
        osp.append(O.OracleSpace(O.HEXA, 2, lv[s.level]["verts"], s.global_numbers(), s.n_dof, s.active_bound,
                                 rp, col, np.full((s.n_cells, 6), -1, np.int32), lv[s.level]["parent"]))
    vals = [M.get_entries() for M in mats]
    bh, xh = b.cpu().numpy(), x0.cpu().numpy()
    t = {}
    for k in iterations:
        op = O.solver_params(iterative_solver_type=0, preconditioner=2, gmres_restart=30, residual_tolerance=0.0,
                             residual_reduction=1e-8, max_n_iterations=k, mg_cycle=0, mg_smoother_coarse=0,
                             mg_n_pre_smooth=2, mg_n_post_smooth=2, jacobi_damp=omega, mg_correction_damp=1.0,
                             mg_coarse_residual=0.1, mg_coarse_max_n_iterations=10)
        t0 = time.perf_counter()
        O.solve(op, osp, vals, bh, xh)
        t[k] = time.perf_counter() - t0
    k0, k1 = iterations
    return {"per_iteration_s": (t[k1] - t[k0]) / (k1 - k0), "runs_s": {str(k): round(v, 2) for k, v in t.items()}}


def fp32_assembly_run(pm, torch, ctx, spaces, mats, rhss, params_ow, solver, b, x0, x64, r0n, A, fs, params):
    """The mixed path with FP32 element matrices (fp32_assembly) on all
    levels: entry error against the FP64 assembly, solve time, residual
    against the FP32- and the FP64-assembled finest matrix, solution against
    the FP64 path's, L2/H1 errors.  Restores the FP64 matrices afterwards."""
    import numpy as np
    v64 = A.get_entries().copy()
    rp = A.get_structure()[0]
    p32 = pm.cd_params("exp_sin_3d", disc="supg", eps=EPS, b=B, hexa_quad_rule=2, overwrite=True,
                       fp32_assembly=True)
    for sp_, M_, f_ in zip(spaces, mats, rhss):
        pm.assemble_cd(ctx, sp_, p32, M_, f_)
    v32 = A.get_entries()
    rmax = np.repeat(np.maximum.reduceat(np.abs(v64), rp[:-1]), np.diff(rp))
    ent = np.abs(v32 - v64) / rmax
    solver.update_matrix(mats, spaces)
    x = x0.clone()
    solver.solve(b, x.clone())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = solver.solve(b, x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    r = b.clone()
    A.apply_scaled_add(x, r, -1.0)
    res32 = torch.linalg.norm(r).item() / r0n
    A.set_entries(v64)
    r = b.clone()
    A.apply_scaled_add(x, r, -1.0)
    res64 = torch.linalg.norm(r).item() / r0n
    err = pm.compute_errors(ctx, fs, params, x)
    out = {"entry_max_rel_error": float(ent.max()), "entry_median_rel_error": float(np.median(ent)),
           "mixed_time_to_solution_s": round(dt, 4), "iterations": rep.iterations,
           "rel_residual_own_matrix": res32, "rel_residual_fp64_matrix": res64,
           "solution_rel_diff_vs_fp64": (torch.max(torch.abs(x - x64)) / torch.max(torch.abs(x64))).item(),
           "l2_h1_errors": [err[0], err[1]]}
    for sp_, M_, f_ in zip(spaces, mats, rhss):
        pm.assemble_cd(ctx, sp_, params_ow, M_, f_)
    solver.update_matrix(mats, spaces)
    return out


def cd2d_config(pm, torch, ctx, with_oracle):
    """CD2D: -lap u = f on the unit square, 64 x 64 Q1 (4,225 DOFs),
    u = sin(pi x) sin(pi y) with Dirichlet data, Gauss 3x3, CG + Jacobi to a
    relative residual of 1e-10 (Iteration_cg.C:22-79); with the oracle's
    solve (one core) beside it in the CPU baseline."""
    prm = pm.cd_params("sine_2d", eps=1.0)
    sp = pm.solver_params(iterative_solver_type="cg", preconditioner="jacobi", residual_tolerance=1e-10,
                          residual_reduction=0.0, max_n_iterations=5000)
    cd = pm.ConvectionDiffusion(ctx, pm.TDomain("UnitSquare", 6), 1, prm, sp)
    cd.assemble()
    cd.solve()
    best = {}
    for _ in range(3):
        cd.solution.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cd.assemble()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rep = cd.solve()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for k, v in (("assembly_s", t1 - t0), ("update_and_solve_s", t2 - t1)):
            best[k] = min(best.get(k, 1e9), v)
    err = cd.output()
    out = {"n_dof": cd.spaces[-1].n_dof, "iterations": rep.iterations, "converged": bool(rep.converged),
           "final_rel_residual": rep.final_residual, "assembly_s": round(best["assembly_s"], 5),
           "update_and_solve_s": round(best["update_and_solve_s"], 5), "l2_error": err[0], "h1_error": err[1]}
    oracle = None
    if with_oracle:
        from oracle import oracle as O
        osp, _ = O.build_space(0, 6, 6, 1)
        op = O.cd_params(example=6, eps=1.0)
        t0 = time.perf_counter()
        val, rhs = O.assemble(osp, op)
        t1 = time.perf_counter()
        import numpy as np
        x0 = np.zeros(osp.n_dof)
        x0[osp.active:] = rhs[osp.active:]
        par = O.solver_params(iterative_solver_type=1, preconditioner=1, residual_tolerance=1e-10,
                              residual_reduction=0.0, max_n_iterations=5000)
        xo, it, _, _ = O.solve(par, [osp], [val], rhs, x0)
        t2 = time.perf_counter()
        eo = O.errors(osp, op, xo)
        oracle = {"iterations": it, "assembly_s": round(t1 - t0, 5), "solve_s": round(t2 - t1, 5),
                  "l2_error": eo[0], "h1_error": eo[1], "cores": 1}
    return out, oracle


def batch_problems(n_problems=256, seed=0x5EED):
    """(n_side, eps, b, c) of the sweep: n ~ U{32..256}, eps log-uniform in
    [1e-4, 1], |b| = 1 at a uniform angle (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = rng.integers(32, 257, n_problems)
    eps = 10.0 ** rng.uniform(-4, 0, n_problems)
    th = rng.uniform(0, 2 * math.pi, n_problems)
    return [(int(n), float(e), (math.cos(t), math.sin(t)), 0.0) for n, e, t in zip(ns, eps, th)]


def batched_sweep(pm, ctx, n_problems=256, steps=20, seed=0x5EED, inner_solver="bi_cgstab"):
    """256 independent TCD2D problems (P1, n ~ U{32..256} squares per side,
    eps log-uniform in [1e-4, 1], |b| = 1 at a uniform angle), backward Euler
    tau = 1e-2 for 20 steps, all advanced together; FP64 residual reduction
    1e-10 per step per problem; FP32 inner BiCGStab (or GMRES(30))."""
    probs = batch_problems(n_problems, seed)
    t0 = time.perf_counter()
    bt = pm.BatchTCD(ctx, probs, tau=1e-2, inner_solver=inner_solver)
    t_setup = time.perf_counter() - t0
    bt.initialize()
    # snapshots of every problem (SnapshotsCollector, steps_per_snap 5: the
    # initial condition and every 5th step), kept in device memory
    per_snap = 5
    bt.reserve_snapshots(steps // per_snap + 1)
    bt.write_snapshot()
    seconds, reps = 0.0, []
    for _ in range(steps // per_snap):
        r = bt.run(per_snap)
        seconds += r.seconds
        reps.append(r)
        bt.write_snapshot()
    rep = reps[-1]
    t0 = time.perf_counter()
    bt.compute_pod(pod_inner_product="L2")
    t_pod = time.perf_counter() - t0
    inf = bt.info()
    inner, outer, failed = bt.iterations()
    return {"problems": n_problems, "steps": steps, "total_dof": inf.total_dof, "total_nnz": inf.total_nnz,
            "seconds": round(seconds, 3), "mdof_steps_per_s": round(inf.total_dof * steps / seconds / 1e6, 2),
            "pod_l2_all_problems_ms": round(t_pod * 1e3, 2), "snapshots_per_problem": bt.n_snaps,
            "inner_iterations_total": int(rep.total_inner_iterations),
            "inner_iterations_max_problem": int(rep.max_inner_iterations),
            "failed_solves": int(rep.failed_solves), "block": int(rep.block),
            "solve_workgroups": int(inf.n_workgroups),
            "host_setup_s": round(t_setup, 2),
            "method": "persistent launch, one workgroup per problem (several for the costliest): "
                      "FP64 residual refinement around FP32 %s + Jacobi" % (
                          "BiCGStab" if inner_solver == "bi_cgstab" else "GMRES(30)")}


def cpu_baseline_batch(batch, n_sample=8, steps=3):
    """The oracle (one core) on a bounded sample of the sweep: backward Euler
    for the first problems' (eps, b) on the 64 x 64 TwoTriangles mesh (4,225
    DOFs, the oracle's meshes are refinements), FGMRES + Jacobi to the same
    relative residual 1e-10 per step; DOF-steps per second and the sweep's
    time at that rate."""
    from threadpoolctl import threadpool_limits
    from oracle import oracle as O
    sp, _ = O.build_space(1, 6, 6, 1)
    prm = O.solver_params(iterative_solver_type=0, preconditioner=1, gmres_restart=30, max_n_iterations=5000,
                          residual_tolerance=0.0, residual_reduction=1e-10)

    def solver(S, rhs, x0):
        x, it, _, _ = O.solve(prm, [sp], [S], rhs, x0)
        return x
    t = 0.0
    with threadpool_limits(1):
        for n, e, bb, c in batch_problems()[:n_sample]:
            t0 = time.perf_counter()
            O.tcd_time_integration(sp, 7, 0, O.BACKWARD_EULER, 1e-2, steps, eps=e, b=bb, c=c, solver=solver)
            t += time.perf_counter() - t0
    rate = sp.n_dof * steps * n_sample / t
    total = batch["total_dof"] * batch["steps"] / rate
    return {"mdof_steps_per_s": round(rate / 1e6, 4), "estimated_sweep_s": round(total, 1), "cores": 1,
            "kind": "port", "sample": "oracle backward Euler, %d problems x %d steps on the 4,225-DOF mesh "
            "(%.1f s incl. assembly and the initial interpolation), FGMRES + Jacobi to 1e-10" % (
                n_sample, steps, t)}


def cpu_baseline_oseen(oseen, iterations=5, nref=3, nu=1e-2, vanka_damp=0.2, pressure="q1"):
    """The oracle (numpy with one BLAS thread) on a bounded sample of the Oseen
    leg: 8^3 Q2/Q1 or Q2/P1disc (2 levels), the same Fourier convection field,
    FGMRES + multigrid V(2,2) with cell Vanka (the leg's damping, correction
    0.8, direct coarse solve) for a few iterations; seconds per iteration per
    DOF times the 32^3 DOFs and the GPU's iteration count."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle import oracle as O
    levels = []
    rhs = None
    for lev in (nref - 1, nref):
        o = O.NseOracle(nref, lev, pressure)
        w = fourier_field(O.nodal_coords(o.vs)).reshape(-1)
        A, B, rhs = o.assemble(nu, w)
        levels.append(O.NseLevelOracle(o, A, B))
    F = levels[-1]
    nuu, na = F.ns.n_u, F.ns.n_u_active
    x0 = np.zeros(F.ns.n)
    for k in range(3):
        x0[k * nuu + na:(k + 1) * nuu] = rhs[k * nuu + na:(k + 1) * nuu]
    with threadpool_limits(1):
        t0 = time.perf_counter()
        O.fgmres(lambda z: F.ns.apply(F.A, F.B, z),
                 lambda v: O.nse_mg_apply(levels, v, damp_corr=0.8, vanka_damp=vanka_damp, n_pre=2, n_post=2),
                 rhs, x0, restart=30, max_it=iterations, tol=0.0)
        t = time.perf_counter() - t0
    per_it_dof = t / iterations / F.ns.n
    it_gpu = oseen["fp64"]["iterations"]
    est = per_it_dof * oseen["n_dof"] * it_gpu
    return {"value": round(est, 1), "unit": "s", "cores": 1, "kind": "port",
            "sample": "oracle FGMRES + cell-Vanka multigrid on 8^3 (%d DOFs, 2 levels), %d iterations in %.2f s: "
                      "per-iteration cost per DOF x %d DOFs x the %d iterations of the GPU FP64 solve" % (
                          F.ns.n, iterations, t, oseen["n_dof"], it_gpu)}


def fourier_field(x, seed=0x5EED, n_modes=8):
    """Sum of n_modes divergence-free Fourier modes a_k cos(2 pi k.x + phi_k),
    a_k orthogonal to k, integer wave vectors in [-2, 2]^3 (seeded)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    w = np.zeros((3, len(x)))
    for _ in range(n_modes):
        k = rng.integers(-2, 3, 3)
        while not k.any():
            k = rng.integers(-2, 3, 3)
        a = rng.standard_normal(3)
        a -= (a @ k) / (k @ k) * k
        w += np.outer(a, np.cos(2 * math.pi * (x @ k) + rng.uniform(0, 2 * math.pi)))
    return w


def oseen_solve(pm, torch, ctx, nref=5, nlev=4, nu=1e-2, vanka_damp=0.2, steps=2, pressure="q1"):
    """Oseen on the unit cube, 32^3 hexahedra, Taylor-Hood Q2/Q1 (or Q2/P1disc
    with pressure="p1disc"; NSTYPE 1), nu = 1e-2, frozen convection = Q2
    interpolant of 8 random divergence-free Fourier modes, rhs and boundary
    data of test_u_2_p_1; FGMRES(30) + 4-level multigrid V(2,2) with the
    cell-Vanka (Jacobi) smoother, correction damping 0.8, direct coarse
    solve, relative residual 1e-8."""
    dom = pm.TDomain("Default_UnitCube_Hexa", nref)
    t0 = time.perf_counter()
    systems = []
    for lev in range(nref - nlev + 1, nref + 1):
        g = pm.NSESystem(ctx, dom, lev, pressure)
        xyz = pm.TFESpace(dom, lev, 2, None).dof_coords()
        w = torch.as_tensor(fourier_field(xyz).reshape(-1)).cuda()
        r = g.zeros()
        g.assemble("u2_p1", nu, w, r)
        systems.append((g, r))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    g, b = systems[-1]
    nuu, na = g.n_u, g.n_u_active

    def start():
        x = g.zeros()
        for k in range(3):
            x[k * nuu + na:(k + 1) * nuu] = b[k * nuu + na:(k + 1) * nuu]
        return x
    r0 = b.clone()
    g.apply_scaled_add(start(), r0, -1.0)
    r0n = torch.linalg.norm(r0).item()
    out = {"pressure": pressure, "n_dof": g.n, "n_u": 3 * g.n_u, "n_p": g.n_p, "levels": nlev,
           "nnz_combined": int(g.info.nnz_combined), "host_and_assembly_setup_s": round(t_setup, 3),
           "vanka_damp": vanka_damp}
    for prec in ("mixed", "fp64"):
        sol = pm.Solver(ctx, pm.solver_params(
            iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30, residual_tolerance=0.0,
            residual_reduction=1e-8, max_n_iterations=400, multigrid_cycle_type="V",
            multigrid_smoother="cell_vanka_jacobi", multigrid_smoother_coarse="direct_solve",
            multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2, multigrid_vanka_damp_factor=vanka_damp,
            multigrid_correction_damp_factor=0.8, precision=prec))
        t0 = time.perf_counter()
        sol.update_saddle([s[0] for s in systems])
        torch.cuda.synchronize()
        t_up = time.perf_counter() - t0
        ts = []
        for _ in range(steps):
            x = start()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rep = sol.solve(b, x)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        r = b.clone()
        g.apply_scaled_add(x, r, -1.0)
        out[prec] = {"time_to_solution_s": round(min(ts), 4), "iterations": rep.iterations,
                     "fp64_rel_residual": torch.linalg.norm(r).item() / r0n, "solver_update_s": round(t_up, 3)}
        del sol
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["time_to_solution_s"] / out["mixed"]["time_to_solution_s"], 3)
    return out


def profile_kernel(pm, torch, ctx, dom, nref, nlev, omega, kernel):
    """A few launches of one workload for rocprofv3 counter runs, then a
    1 GiB copy (torch copy_) three times as the calibration of FETCH_SIZE /
    WRITE_SIZE.  kernel: jacobi_bf16 / apply_fp64 / ... (the FE matrix's own
    SELL kernels), solver_jacobi_bf16 / solver_jacobi_fp64 (the multigrid's
    finest-level sweep), assembly (the 64^3 Q2 SUPG assembly, 6 times per
    route: element GEMM FP64 and FP32, quadrature kernel),
    vanka / vanka_p1disc (Oseen 32^3 Q2/Q1 or Q2/P1disc: 6 mixed-precision
    FGMRES iterations with the cell-Vanka multigrid), batch / batch_gmres (the
    256-problem TCD2D sweep with BiCGStab or GMRES, 2 steps)."""
    params = pm.cd_params("exp_sin_3d", disc="supg", eps=EPS, b=B, hexa_quad_rule=2)
    n_rows, nnz, ms = 0, 0, 0.0
    if kernel == "assembly":
        s = pm.TFESpace(dom, nref, 2, ctx)
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        for kw in (dict(overwrite=True), dict(overwrite=True, fp32_assembly=True), dict(assembly_variant=1)):
            prm = pm.cd_params("exp_sin_3d", disc="supg", eps=EPS, b=B, hexa_quad_rule=2, **kw)
            ts = []
            for _ in range(6):
                A.reset()
                f.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                pm.assemble_cd(ctx, s, prm, A, f)
                ts.append(time.perf_counter() - t0)
            print("assembly", kw, "%.3f ms" % (1e3 * min(ts[1:])), flush=True)
        n_rows, nnz, ms = A.n_rows, A.nnz, 1e3 * min(ts[1:])
    elif kernel in ("vanka", "vanka_p1disc"):
        p1 = kernel == "vanka_p1disc"
        dom5 = pm.TDomain("Default_UnitCube_Hexa", 5)
        systems = []
        for lev in range(2, 6):
            g = pm.NSESystem(ctx, dom5, lev, "p1disc" if p1 else "q1")
            xyz = pm.TFESpace(dom5, lev, 2, None).dof_coords()
            w = torch.as_tensor(fourier_field(xyz).reshape(-1)).cuda()
            r = g.zeros()
            g.assemble("u2_p1", 1e-2, w, r)
            systems.append((g, r))
        sol = pm.Solver(ctx, pm.solver_params(
            iterative_solver_type="fgmres", preconditioner="multigrid", gmres_restart=30, residual_tolerance=0.0,
            residual_reduction=1e-30, max_n_iterations=6, multigrid_cycle_type="V",
            multigrid_smoother="cell_vanka_jacobi", multigrid_smoother_coarse="direct_solve",
            multigrid_n_pre_smooth=2, multigrid_n_post_smooth=2, multigrid_vanka_damp_factor=0.8 if p1 else 0.2,
            multigrid_correction_damp_factor=0.8, precision="mixed"))
        sol.update_saddle([q[0] for q in systems])
        g, b = systems[-1]
        x = g.zeros()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol.solve(b, x)
        torch.cuda.synchronize()
        n_rows, nnz, ms = g.n, int(g.info.nnz_combined), 1e3 * (time.perf_counter() - t0)
    elif kernel in ("batch", "batch_gmres"):
        import numpy as np
        rng = np.random.default_rng(0x5EED)
        ns = rng.integers(32, 257, 256)
        eps = 10.0 ** rng.uniform(-4, 0, 256)
        th = rng.uniform(0, 2 * math.pi, 256)
        bt = pm.BatchTCD(ctx, [(int(n_), float(e), (math.cos(t), math.sin(t)), 0.0) for n_, e, t in zip(ns, eps, th)],
                         tau=1e-2, inner_solver="gmres" if kernel == "batch_gmres" else "bi_cgstab")
        bt.initialize()
        r = bt.run(2)
        inf = bt.info()
        n_rows, nnz, ms = inf.total_dof, inf.total_nnz, 1e3 * r.seconds / 2
    elif kernel.startswith("solver_"):
        # the finest-level sweep as the multigrid runs it (solver-internal order)
        spaces = [pm.TFESpace(dom, l, 2, ctx) for l in range(nref - nlev + 1, nref + 1)]
        mats = []
        for sp in spaces:
            M = pm.FEMatrix(ctx, sp)
            f = torch.zeros(sp.n_dof, dtype=torch.float64, device="cuda")
            pm.assemble_cd(ctx, sp, params, M, f)
            mats.append(M)
        k = kernel[len("solver_"):]
        sol = pm.Solver(ctx, solver_params(pm, "mixed" if k.endswith("bf16") else "fp64", omega))
        sol.update_matrix(mats, spaces)
        ms = sol.time_level_kernel(nlev - 1, k, reps=5)
        A = mats[-1]
    else:
        s = pm.TFESpace(dom, nref, 2, ctx)
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.assemble_cd(ctx, s, params, A, f)
        ms = A.time_kernel(kernel, reps=5)
    if kernel.startswith("solver_") or kernel not in ("assembly", "vanka", "vanka_p1disc", "batch", "batch_gmres"):
        n_rows, nnz = A.n_rows, A.nnz
    src = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kernel": kernel, "n_rows": n_rows, "nnz": nnz, "ms": ms,
                      "copy_bytes": 2 * src.numel() * 8}))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of this build has a single MI355X and the "
              "solve path has no multi-GPU leg (domain decomposition is described, not built).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from mymoonmd_amd import parmoon as pm

    nref, nlev = args.refine, args.levels
    ctx = pm.Context(0)
    dom = pm.TDomain("Default_UnitCube_Hexa", nref)
    if args.profile_kernel:
        profile_kernel(pm, torch, ctx, dom, nref, nlev, args.jacobi_damp, args.profile_kernel)
        return
    levels = list(range(nref - nlev + 1, nref + 1))
    t0 = time.perf_counter()
    spaces = [pm.TFESpace(dom, l, 2, ctx) for l in levels]
    t_setup_spaces = time.perf_counter() - t0
    params = pm.cd_params("exp_sin_3d", disc="supg", eps=EPS, b=B, hexa_quad_rule=2)
    mats, rhss = [], []
    for s in spaces:
        A = pm.FEMatrix(ctx, s)
        f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
        pm.assemble_cd(ctx, s, params, A, f)
        mats.append(A)
        rhss.append(f)
    fs, A, b = spaces[-1], mats[-1], rhss[-1]
    ab = fs.active_bound

    # assembly + one SpMV (throughput in MDoF/s), inputs resident.  The
    # assembly sets matrix and rhs (overwrite: the reference's reset + add in
    # one pass); the same for each assembly route alone
    params_ow = pm.cd_params("exp_sin_3d", disc="supg", eps=EPS, b=B, hexa_quad_rule=2, overwrite=True)
    x1 = torch.ones(fs.n_dof, dtype=torch.float64, device="cuda")
    y1 = torch.empty_like(x1)
    f_tmp = torch.zeros_like(b)
    t_as = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pm.assemble_cd(ctx, fs, params_ow, A, f_tmp)
        A.apply(x1, y1)
        t_as.append(time.perf_counter() - t0)
    t_asm_spmv = min(t_as)
    t_routes = {}
    for name, kw in (("element_gemm_fp64", {}), ("element_gemm_fp32", dict(fp32_assembly=True)),
                     ("quadrature_kernel", dict(assembly_variant=1))):
        prm = pm.cd_params("exp_sin_3d", disc="supg", eps=EPS, b=B, hexa_quad_rule=2, overwrite=True, **kw)
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pm.assemble_cd(ctx, fs, prm, A, f_tmp)
            ts.append(time.perf_counter() - t0)
        t_routes[name] = round(1000 * min(ts), 3)
    pm.assemble_cd(ctx, fs, params_ow, A, f_tmp)
    del f_tmp, x1, y1

    solvers = {}
    t_update = {}
    for prec in ("mixed", "fp64"):
        t0 = time.perf_counter()
        sol = pm.Solver(ctx, solver_params(pm, prec, args.jacobi_damp, renumber=args.renumber_levels,
                                           inner=args.inner_reduction))
        sol.update_matrix(mats, spaces)
        t_update[prec] = time.perf_counter() - t0
        solvers[prec] = sol

    x0 = torch.zeros_like(b)
    x0[ab:] = b[ab:]
    r0 = b.clone()
    A.apply_scaled_add(x0, r0, -1.0)
    r0n = torch.linalg.norm(r0).item()

    def run(prec):
        x = x0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = solvers[prec].solve(b, x)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        r = b.clone()
        A.apply_scaled_add(x, r, -1.0)
        rel = torch.linalg.norm(r).item() / r0n
        return dt, rep.iterations, rel, x

    for _ in range(args.warmup):
        run("mixed")
        run("fp64")
    times = {"mixed": [], "fp64": []}
    its, rels = {}, {}
    xs = {}
    torch.cuda.synchronize()
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        for prec in ("mixed", "fp64"):
            dt, it, rel, x = run(prec)
            times[prec].append(dt)
            its[prec], rels[prec], xs[prec] = it, rel, x
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t_all0
    tts = {k: sum(v) / len(v) for k, v in times.items()}
    # the finest-level smoothing sweep as the cycles ran it (renumbered copy)
    kt_solver = {"jacobi_bf16": solvers["mixed"].time_level_kernel(nlev - 1, "jacobi_bf16", reps=20),
                 "jacobi_fp64": solvers["fp64"].time_level_kernel(nlev - 1, "jacobi_fp64", reps=20)}
    sol_diff = (torch.max(torch.abs(xs["mixed"] - xs["fp64"])) /
                torch.max(torch.abs(xs["fp64"]))).item()
    errs = {k: pm.compute_errors(ctx, fs, params, xs[k]) for k in xs}

    # the error of each stage of the mixed path against the FP64 path (whose
    # assembly, operator and cycle match the oracle at 1e-12 / 1e-10 in the
    # tests): one multigrid cycle on b, the solution, its FP64 residual
    zs = {}
    for prec in ("mixed", "fp64"):
        z = torch.zeros_like(b)
        solvers[prec].apply_preconditioner(b.clone(), z)
        zs[prec] = z
    stage_errors = {
        "assembly": "FP64 in both paths (the same kernel); FP32 element matrices: fp32_assembly",
        "bf16_matrix_max_rel_entry_error": bf16_row_scaled_error(A.get_entries(), A.get_structure()[0]),
        "cycle_rel_diff": (torch.max(torch.abs(zs["mixed"] - zs["fp64"])) / torch.max(torch.abs(zs["fp64"]))).item(),
        "solution_rel_diff": sol_diff,
        "fp64_verified_rel_residual": rels["mixed"]}
    del zs

    # Solver::update_matrix with the same hierarchy (values only), and the
    # whole step as a time loop or a Picard iteration sees it: assembly of
    # all levels + value update + solve
    t_refresh, t_e2e = {}, {}
    for prec in ("mixed", "fp64"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solvers[prec].update_matrix(mats, spaces)
        torch.cuda.synchronize()
        t_refresh[prec] = time.perf_counter() - t0
        ts = []
        for _ in range(2):
            x = x0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for sp_, M_, f_ in zip(spaces, mats, rhss):
                pm.assemble_cd(ctx, sp_, params_ow, M_, f_)
            solvers[prec].update_matrix(mats, spaces)
            solvers[prec].solve(b, x)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        t_e2e[prec] = min(ts)

    # FP32 element matrices (FP32 matrix cores) on every level, then the
    # mixed solve: its residual against the FP32-assembled matrix and against
    # the FP64 one, the solution against the FP64 path's, the errors
    fp32asm = fp32_assembly_run(pm, torch, ctx, spaces, mats, rhss, params_ow, solvers["mixed"], b, x0,
                                xs["fp64"], r0n, A, fs, params)

    # configuration 3's tolerance: 1e-10 (mixed and FP64)
    tight = {}
    for prec in ("mixed", "fp64"):
        sol10 = pm.Solver(ctx, solver_params(pm, prec, args.jacobi_damp, renumber=args.renumber_levels,
                                             inner=args.inner_reduction, reduction=1e-10))
        sol10.update_matrix(mats, spaces)
        x = x0.clone()
        sol10.solve(b, x.clone())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep10 = sol10.solve(b, x)
        torch.cuda.synchronize()
        dt10 = time.perf_counter() - t0
        r = b.clone()
        A.apply_scaled_add(x, r, -1.0)
        tight[prec] = {"time_to_solution_s": round(dt10, 4), "iterations": rep10.iterations,
                       "fp64_verified_rel_residual": torch.linalg.norm(r).item() / r0n}
        del sol10
    tight["speedup_mixed_vs_fp64"] = round(tight["fp64"]["time_to_solution_s"] /
                                           tight["mixed"]["time_to_solution_s"], 3)

    # the configured damping 0.8, once, with an iteration cap (outside the timed region)
    damp08 = None
    if args.damp08_iterations > 0:
        del solvers
        damp08 = {}
        for prec in ("mixed", "fp64"):
            sol = pm.Solver(ctx, solver_params(pm, prec, 0.8, args.damp08_iterations))
            sol.update_matrix(mats, spaces)
            x = x0.clone()
            t0 = time.perf_counter()
            rep = sol.solve(b, x)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            r = b.clone()
            A.apply_scaled_add(x, r, -1.0)
            damp08[prec] = {"iterations": rep.iterations, "converged": bool(rep.converged),
                            "fp64_rel_residual": torch.linalg.norm(r).item() / r0n, "time_s": round(dt, 3)}
            del sol

    # roofline of the dominant kernel: the fused bf16 Jacobi sweep of the
    # finest level (3 full sweeps + the defect of the 5 finest-level operator
    # passes of a V(2,2) cycle; the first sweep from zero reads no matrix), as
    # the solver runs it on its renumbered copy
    nnz, n = A.nnz, A.n_rows
    kt = {k: A.time_kernel(k, reps=20) for k in ("apply_fp64", "jacobi_fp64", "apply_bf16", "jacobi_bf16")}
    bytes_alg = {"apply_fp64": 12 * nnz + 20 * n, "jacobi_fp64": 12 * nnz + 20 * n + 24 * n,
                 "apply_bf16": 6 * nnz + 16 * n, "jacobi_bf16": 6 * nnz + 16 * n + 12 * n}
    # the solver's sweep reads the finest level's folded copy with slices of
    # equal offset patterns: per entry of an active row in an active column
    # only its value (bf16: 2 B, FP64: 8 B; the offsets are read once per
    # slice); per active row the row map, x, b, D^-1 and the written x' (bf16:
    # 4 B each + the FP32 row scale; FP64: 8 B each) and, outside the
    # refinement, the Dirichlet-column sum; per Dirichlet row x, b, D^-1, x'.
    # Inside the refinement (mixed) the vectors vanish on the Dirichlet rows
    # and the sums are not read.
    n_act, nnz_aa = folded_entries(A, fs.active_bound)
    n_slices = (n_act + 63) // 64
    bytes_solver = {"jacobi_bf16": 2 * nnz_aa + 24 * n_act + 16 * (n - n_act) + 12 * n_slices,
                    "jacobi_fp64": 8 * nnz_aa + 44 * n_act + 32 * (n - n_act) + 12 * n_slices}
    dom_kernel = "jacobi_bf16"
    ach = bytes_solver[dom_kernel] / (kt_solver[dom_kernel] * 1e-3) / 1e9
    roof = {"bound": "hbm",
            "kernel": "k_sell_pat<MODE_JACOBI, bf16, float, GSUM=false> (finest level, solver-internal "
                      "lexicographic order, slices of equal offset patterns, as the refinement's cycles run it)",
            "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic("solver_pat_" + dom_kernel),
            "bytes_per_launch": bytes_solver[dom_kernel], "avg_launch_ms": round(kt_solver[dom_kernel], 4)}
    spmv_fp64 = bytes_alg["apply_fp64"] / (kt["apply_fp64"] * 1e-3) / 1e9

    cd2d, cd2d_cpu = cd2d_config(pm, torch, ctx, not args.no_cpu_baseline)
    batch = None if args.no_batch else batched_sweep(pm, ctx)
    batch_gmres = None if args.no_batch else batched_sweep(pm, ctx, steps=20, inner_solver="gmres")
    oseen = None if args.no_oseen else oseen_solve(pm, torch, ctx)
    # the reference's cell-Vanka pair (Q2/P1disc) at the same size, with the
    # damping of its own Vanka test (0.8)
    oseen_p1 = None if args.no_oseen else oseen_solve(pm, torch, ctx, vanka_damp=0.8, pressure="p1disc")

    cpu = None
    if not args.no_cpu_baseline:
        nl_cpu = max(2, args.cpu_refine - 1)
        c = cpu_baseline(args.cpu_refine, nl_cpu, args.jacobi_damp)
        full = cpu_baseline_full(pm, nref, spaces, mats, b, x0, args.jacobi_damp)
        est = full["per_iteration_s"] * its["fp64"]
        cpu = {"value": round(est, 2), "unit": "s", "cores": 1, "kind": "port", "cpu_model": cpu_model(),
               "sample": "oracle (sequential C restatement, one core) FGMRES(30)+MG V(2,2) on the full 64^3 "
                         "problem (%d DOFs, 4 levels, coarse level 10 Jacobi steps), runs of 1 and 2 iterations "
                         "(%s s): %.2f s per full-size iteration x the %d iterations of the GPU FP64 solve. "
                         "Oracle assembly+SpMV at 16^3: %.3f MDoF/s (GPU/CPU %.0fx)" % (
                             n, full["runs_s"], full["per_iteration_s"], its["fp64"], c["mdofs"],
                             (n / t_asm_spmv / 1e6) / c["mdofs"]),
               "cd2d": cd2d_cpu}
        if batch is not None:
            cpu["batched_tcd2d"] = cpu_baseline_batch(batch)
        if oseen is not None:
            cpu["oseen_q2q1_32"] = cpu_baseline_oseen(oseen)
            cpu["oseen_q2p1disc_32"] = cpu_baseline_oseen(oseen_p1, vanka_damp=0.8, pressure="p1disc")

    out = {
        "metric": "time-to-solution @1e-8 rel. residual, mixed vs FP64 GPU; assembly+SpMV MDoF/s",
        "value": round(tts["mixed"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * tts["mixed"], 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 outer / f32+bf16 V-cycle",
        "data": "synthetic (analytic exp-sin solution, seeded nothing random)",
        "config": {"workload": "CD3D SUPG Q2 hexa %d^3 (%d DOF, %d nnz), FGMRES(30)+MG V(2,2) "
                               "Jacobi %.1f, %d levels, 1e-8 rel. residual" % (
                                   2 ** nref, n, nnz, args.jacobi_damp, nlev)},
        "time_to_solution_s": {"mixed": round(tts["mixed"], 4), "fp64": round(tts["fp64"], 4)},
        "speedup_mixed_vs_fp64": round(tts["fp64"] / tts["mixed"], 3),
        "iterations": its,
        "fp64_verified_rel_residual": rels,
        "mixed_vs_fp64_solution_rel_diff": sol_diff,
        "l2_h1_errors": {k: [errs[k][0], errs[k][1]] for k in errs},
        "assembly_spmv_mdofs": round(n / t_asm_spmv / 1e6, 2),
        "assembly_spmv_ms": round(1000 * t_asm_spmv, 3),
        "assembly_ms_by_route": t_routes,
        "fp32_assembly": fp32asm,
        "solver_update_s": {k: round(v, 3) for k, v in t_update.items()},
        "solver_value_update_s": {k: round(v, 4) for k, v in t_refresh.items()},
        "assembly_update_solve_s": {k: round(v, 4) for k, v in t_e2e.items()},
        "mixed_stage_errors": stage_errors,
        "reduction_1e-10": tight,
        "kernel_ms_reference_order": {k: round(v, 4) for k, v in kt.items()},
        "kernel_ms_solver_finest": {k: round(v, 4) for k, v in kt_solver.items()},
        "spmv_fp64_gbs": round(spmv_fp64, 1),
        "timed_region_s": round(t_all, 3),
        "jacobi_damp_0.8_check": damp08,
        "cd2d_poisson_64": cd2d,
        "batched_tcd2d": batch,
        "batched_tcd2d_gmres": batch_gmres,
        "oseen_q2q1_32": oseen,
        "oseen_q2p1disc_32": oseen_p1,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
