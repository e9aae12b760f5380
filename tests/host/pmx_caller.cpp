// A ParMooN-side caller of the C ABI, written against include/pmx.h alone
// (plus the HIP runtime for device buffers), in the shape of the adapter a
// maintainer adds to ConvectionDiffusion<3> (src/System/ConvectionDiffusion.C:
// 310-345 assemble, Solver<>::update_matrix + solve): a CD3D problem on the
// unit cube, three multigrid levels of Q2 hexahedra.
//
// What it does, each step checked, then one JSON line on stdout:
//  1. the coefficients of a CoeffFct3D (here the exp-sin example written out
//     in C++) sampled at pmx_assembly_quad_points and handed over as
//     pmx_cd_params.dev_coeff_table, the boundary values evaluated at
//     pmx_space_get_dof_coords as host_dirichlet: the finest level's matrix
//     equals the built-in example's on the same (quadrature-kernel) route
//     bit for bit, the load vector within 1e-13, and the default
//     element-GEMM route (with overwrite, on unzeroed buffers) within 1e-13;
//  2. FGMRES(30) + multigrid V(2,2) (mixed precision, direct coarse solve)
//     to a relative residual of 1e-8, the residual verified with
//     pmx_vector_copy + pmx_matrix_apply_scaled_add + pmx_vector_norm and
//     once more on the host from the CSR arrays;
//  3. the L2 / H1 errors of the solution (pmx_compute_errors);
//  4. an invalid call returns PMX_E_ARG with a message, as the adapter's
//     ErrThrow path expects.
// Exit status 0 only when every check holds.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "pmx.h"

namespace {

void check(int rc, const char* what) {
  if (rc != PMX_OK) throw std::runtime_error(std::string(what) + ": " + pmx_last_error());
}
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T* dev_alloc(size_t n) {
  void* p = nullptr;
  hip_check(hipMalloc(&p, n * sizeof(T) + 1), "hipMalloc");
  return static_cast<T*>(p);
}

// the example's data as a user's CoeffFct3D / BoundCondition would give them
const double kPi = 3.14159265358979323846;
double exact(double x, double y, double z) { return std::exp(-x) * std::sin(kPi * y) * std::sin(kPi * z); }
void coeffs(double x, double y, double z, double eps, const double* b, double c, double* out8) {
  const double u = exact(x, y, z);
  const double ux = -u;
  const double uy = kPi * std::exp(-x) * std::cos(kPi * y) * std::sin(kPi * z);
  const double uz = kPi * std::exp(-x) * std::sin(kPi * y) * std::cos(kPi * z);
  const double lap = (1.0 - 2.0 * kPi * kPi) * u;
  out8[0] = eps;
  out8[1] = b[0];
  out8[2] = b[1];
  out8[3] = b[2];
  out8[4] = c;
  out8[5] = -eps * lap + b[0] * ux + b[1] * uy + b[2] * uz + c * u;
  out8[6] = out8[7] = 0.0;
}

struct Level {
  pmx_space* space = nullptr;
  pmx_matrix* A = nullptr;
  pmx_space_info info{};
  double* rhs = nullptr;
  double* table = nullptr;
  std::vector<double> dirichlet;
};

}  // namespace

int main(int argc, char** argv) {
  const int n_refine = argc > 1 ? std::atoi(argv[1]) : 3;
  const int n_levels = 3;
  pmx_ctx* ctx = nullptr;
  pmx_domain* dom = nullptr;
  pmx_solver* solver = nullptr;
  std::vector<Level> lv(n_levels);
  int failures = 0;
  try {
    check(pmx_ctx_create(0, &ctx), "pmx_ctx_create");
    check(pmx_domain_create(PMX_GEO_UNIT_CUBE_HEXA, n_refine, &dom), "pmx_domain_create");
    const double eps = 1e-3, c = 0.0;
    const double b[3] = {1 / std::sqrt(14.0), 2 / std::sqrt(14.0), 3 / std::sqrt(14.0)};
    const int rule = 2;   // INTERNAL_QUAD_RULE = 2: Gauss3 on hexahedra

    pmx_cd_params base{};
    base.disc = PMX_DISC_SUPG;
    base.example = PMX_EX_EXP_SIN_3D;
    base.hexa_quad_rule = rule;
    base.eps = eps;
    for (int d = 0; d < 3; ++d) base.b[d] = b[d];
    base.c = c;
    base.delta0 = 1.0;
    base.poly_degree = 1.0;
    base.theta1 = 1.0;

    for (int l = 0; l < n_levels; ++l) {
      Level& L = lv[l];
      const int level = n_refine - n_levels + 1 + l;
      check(pmx_space_create(ctx, dom, level, 2, nullptr, nullptr, nullptr, PMX_BC_ALL_DIRICHLET, &L.space),
            "pmx_space_create");
      check(pmx_space_get_info(L.space, &L.info), "pmx_space_get_info");
      check(pmx_matrix_create(ctx, L.space, &L.A), "pmx_matrix_create");
      // coefficient table at the assembly's quadrature points
      int nq = 0;
      check(pmx_assembly_n_quad(L.space, rule, &nq), "pmx_assembly_n_quad");
      std::vector<double> xyz((size_t)L.info.n_cells * nq * 3), tab((size_t)L.info.n_cells * nq * 8);
      check(pmx_assembly_quad_points(L.space, rule, xyz.data()), "pmx_assembly_quad_points");
      for (size_t k = 0; k < (size_t)L.info.n_cells * nq; ++k)
        coeffs(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], eps, b, c, &tab[8 * k]);
      L.table = dev_alloc<double>(tab.size());
      hip_check(hipMemcpy(L.table, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice), "copy table");
      // boundary values at the nodal points of the Dirichlet DOFs
      std::vector<double> dof((size_t)L.info.n_dof * 3);
      check(pmx_space_get_dof_coords(L.space, dof.data()), "pmx_space_get_dof_coords");
      for (int i = L.info.active_bound; i < L.info.n_dof; ++i)
        L.dirichlet.push_back(exact(dof[3 * i], dof[3 * i + 1], dof[3 * i + 2]));
      L.rhs = dev_alloc<double>(L.info.n_dof);
      hip_check(hipMemset(L.rhs, 0, L.info.n_dof * sizeof(double)), "hipMemset");
      pmx_cd_params p = base;
      p.dev_coeff_table = L.table;
      p.host_dirichlet = L.dirichlet.data();
      check(pmx_matrix_reset(ctx, L.A), "pmx_matrix_reset");
      check(pmx_assemble_cd(ctx, L.space, &p, L.A, L.rhs), "pmx_assemble_cd (table)");
    }

    // 1. the same finest level through the built-in example: on the route
    // the table takes (the quadrature kernel) bit for bit, and on the
    // default element-GEMM route (constant coefficients) to rounding
    Level& F = lv.back();
    const int n = F.info.n_dof;
    pmx_matrix* A0 = nullptr;
    pmx_matrix* A2 = nullptr;
    check(pmx_matrix_create(ctx, F.space, &A0), "pmx_matrix_create");
    check(pmx_matrix_create(ctx, F.space, &A2), "pmx_matrix_create");
    double* rhs0 = dev_alloc<double>(n);
    double* rhs2 = dev_alloc<double>(n);
    hip_check(hipMemset(rhs0, 0, n * sizeof(double)), "hipMemset");
    pmx_cd_params pq = base;
    pq.assembly_variant = 1;
    check(pmx_assemble_cd(ctx, F.space, &pq, A0, rhs0), "pmx_assemble_cd (example)");
    pmx_cd_params pg = base;
    pg.overwrite = 1;   // A2 and rhs2 are not zeroed: overwrite sets them
    check(pmx_assemble_cd(ctx, F.space, &pg, A2, rhs2), "pmx_assemble_cd (example, element GEMM)");
    // the library works on its own non-blocking stream: wait for it before
    // reading its results with plain HIP copies
    check(pmx_ctx_synchronize(ctx), "pmx_ctx_synchronize");
    int nr = 0, nnz = 0, ab = 0;
    check(pmx_matrix_get_info(F.A, &nr, &nnz, &ab), "pmx_matrix_get_info");
    std::vector<double> v1(nnz), v0(nnz), v2(nnz), f1(n), f0(n), f2(n);
    check(pmx_matrix_get_entries(ctx, F.A, v1.data()), "pmx_matrix_get_entries");
    check(pmx_matrix_get_entries(ctx, A0, v0.data()), "pmx_matrix_get_entries");
    check(pmx_matrix_get_entries(ctx, A2, v2.data()), "pmx_matrix_get_entries");
    hip_check(hipMemcpy(f1.data(), F.rhs, n * sizeof(double), hipMemcpyDeviceToHost), "copy rhs");
    hip_check(hipMemcpy(f0.data(), rhs0, n * sizeof(double), hipMemcpyDeviceToHost), "copy rhs");
    hip_check(hipMemcpy(f2.data(), rhs2, n * sizeof(double), hipMemcpyDeviceToHost), "copy rhs");
    const bool same_matrix = std::memcmp(v1.data(), v0.data(), nnz * sizeof(double)) == 0;
    double fmax = 0, fdiff = 0, vmax = 0, vdiff = 0, gdiff = 0;
    for (int i = 0; i < n; ++i) {
      fmax = std::fmax(fmax, std::fabs(f0[i]));
      fdiff = std::fmax(fdiff, std::fabs(f1[i] - f0[i]));
      gdiff = std::fmax(gdiff, std::fabs(f2[i] - f0[i]));
    }
    for (int k = 0; k < nnz; ++k) {
      vmax = std::fmax(vmax, std::fabs(v0[k]));
      vdiff = std::fmax(vdiff, std::fabs(v2[k] - v0[k]));
    }
    const double rhs_rel = fdiff / fmax;
    const double gemm_rel = std::fmax(vdiff / vmax, gdiff / fmax);
    if (!same_matrix) ++failures;
    if (!(rhs_rel <= 1e-13)) ++failures;
    if (!(gemm_rel <= 1e-13)) ++failures;
    check(pmx_matrix_destroy(A0), "pmx_matrix_destroy");
    check(pmx_matrix_destroy(A2), "pmx_matrix_destroy");
    hip_check(hipFree(rhs0), "hipFree");
    hip_check(hipFree(rhs2), "hipFree");

    // 2. solver: FGMRES(30) + V(2,2), jacobi 0.7, direct coarse solve, mixed
    pmx_solver_params sp;
    check(pmx_solver_default_params(&sp), "pmx_solver_default_params");
    sp.iterative_solver_type = PMX_IT_FGMRES;
    sp.preconditioner = PMX_PC_MULTIGRID;
    sp.gmres_restart = 30;
    sp.max_n_iterations = 200;
    sp.residual_tolerance = 0.0;
    sp.residual_reduction = 1e-8;
    sp.mg_cycle = PMX_CYCLE_V;
    sp.mg_smoother = PMX_SMOOTHER_JACOBI;
    sp.mg_smoother_coarse = PMX_SMOOTHER_DIRECT;
    sp.mg_n_pre_smooth = 2;
    sp.mg_n_post_smooth = 2;
    sp.jacobi_damp = 0.7;
    sp.precision = PMX_PRECISION_MIXED;
    check(pmx_solver_create(ctx, &sp, &solver), "pmx_solver_create");
    std::vector<pmx_matrix*> mats;
    std::vector<const pmx_space*> spaces;
    for (auto& L : lv) {
      mats.push_back(L.A);
      spaces.push_back(L.space);
    }
    check(pmx_solver_update(solver, mats.data(), spaces.data(), n_levels), "pmx_solver_update");
    double* x = dev_alloc<double>(n);
    std::vector<double> x0(n, 0.0);
    for (int i = F.info.active_bound; i < n; ++i) x0[i] = f1[i];   // solution.copy_nonactive(rhs)
    hip_check(hipMemcpy(x, x0.data(), n * sizeof(double), hipMemcpyHostToDevice), "copy x0");
    double* r = dev_alloc<double>(n);
    check(pmx_vector_copy(ctx, n, F.rhs, r), "pmx_vector_copy");   // ordered on the library's stream
    check(pmx_matrix_apply_scaled_add(ctx, F.A, x, r, -1.0), "pmx_matrix_apply_scaled_add");
    double r0 = 0;
    check(pmx_vector_norm(ctx, n, r, &r0), "pmx_vector_norm");
    pmx_solve_report rep{};
    check(pmx_solver_solve(solver, F.rhs, x, &rep), "pmx_solver_solve");
    check(pmx_vector_copy(ctx, n, F.rhs, r), "pmx_vector_copy");   // ordered on the library's stream
    check(pmx_matrix_apply_scaled_add(ctx, F.A, x, r, -1.0), "pmx_matrix_apply_scaled_add");
    double r1 = 0;
    check(pmx_vector_norm(ctx, n, r, &r1), "pmx_vector_norm");
    const double rel = r1 / r0;
    // the same residual on the host from the CSR arrays
    std::vector<int> rp(n + 1), col(nnz);
    check(pmx_matrix_get_structure(F.A, rp.data(), col.data()), "pmx_matrix_get_structure");
    std::vector<double> xh(n);
    hip_check(hipMemcpy(xh.data(), x, n * sizeof(double), hipMemcpyDeviceToHost), "copy x");
    double h0 = 0, h1 = 0;
    for (int i = 0; i < n; ++i) {
      double ax = 0, a0 = 0;
      if (i < ab) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
          ax += v1[k] * xh[col[k]];
          a0 += v1[k] * x0[col[k]];
        }
      } else {
        ax = xh[i];
        a0 = x0[i];
      }
      h1 += (f1[i] - ax) * (f1[i] - ax);
      h0 += (f1[i] - a0) * (f1[i] - a0);
    }
    const double host_rel = std::sqrt(h1 / h0);
    if (!rep.converged || !(rel <= 1.01e-8) || !(host_rel <= 1.01e-8)) ++failures;

    // 3. errors against the example's exact solution
    double err[4] = {0, 0, 0, 0};
    check(pmx_compute_errors(ctx, F.space, &base, x, err), "pmx_compute_errors");
    if (!(err[0] > 0 && err[0] < 1e-2)) ++failures;

    // 4. the error path of the adapter
    const int bad = pmx_solver_solve(solver, nullptr, x, &rep);
    const std::string msg = pmx_last_error();
    if (bad != PMX_E_ARG || msg.empty()) ++failures;

    std::printf("{\"n_dof\": %d, \"nnz\": %d, \"matrix_equal\": %s, \"rhs_rel_diff\": %.3e, \"gemm_route_rel_diff\": %.3e, \"iterations\": %d, "
                "\"fp64_rel_residual\": %.6e, \"host_rel_residual\": %.6e, \"l2\": %.15e, \"h1\": %.15e, \"error_code\": %d, "
                "\"error_message\": \"%s\", \"failures\": %d}\n",
                n, nnz, same_matrix ? "true" : "false", rhs_rel, gemm_rel, rep.iterations, rel, host_rel, err[0], err[1], bad,
                msg.c_str(), failures);
    hip_check(hipFree(x), "hipFree");
    hip_check(hipFree(r), "hipFree");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "pmx_caller: %s\n", e.what());
    failures = 1000;
  }
  if (solver) pmx_solver_destroy(solver);
  for (auto& L : lv) {
    if (L.A) pmx_matrix_destroy(L.A);
    if (L.space) pmx_space_destroy(L.space);
    if (L.table) (void)hipFree(L.table);
    if (L.rhs) (void)hipFree(L.rhs);
  }
  if (dom) pmx_domain_destroy(dom);
  if (ctx) pmx_ctx_destroy(ctx);
  return failures == 0 ? 0 : 1;
}
