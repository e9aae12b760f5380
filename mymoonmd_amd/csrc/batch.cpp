// Batched time-dependent convection-diffusion: many independent P1 problems on
// unit-square triangulations of different sizes and coefficients, advanced by
// backward Euler together (C ABI pmx_batch_* of include/pmx.h).
//
// Per problem the result is that of TimeConvectionDiffusion<2> with the
// Galerkin TCD forms and backward Euler (src/System/TimeConvectionDiffusion.C:
// 222-386, src/TimeDiscRoutines/TimeDiscretizations.C:106-210).  The
// coefficients are time-independent, so the stiffness and mass matrices the
// reference re-assembles every step are assembled once: the values are the
// same.  The load vector is assembled at every step.
//
// Device layout: see kernels_batch.hip.  One time step is a fixed sequence of
// kernels (advance time, load vector at t, Dirichlet values, rhs_td, the
// persistent solve) captured once into a HIP graph and replayed per step; the
// time lives in device memory so that the replays see the new time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

#include "hip/launch.hpp"
#include "host/fe.hpp"
#include "host/mesh.hpp"
#include "host/space.hpp"
#include "internal.hpp"
#include "pmx.h"

using namespace pmx_internal;

struct pmx_batch {
  CtxRef ctx;
  pmx_batch_params prm{};
  int n_prob = 0;
  std::vector<pmx_batch_problem> probs;
  std::vector<pmx::Mesh> meshes;
  std::vector<pmx::FESpace> spaces;
  std::vector<pmx::CSRPattern> pats;
  std::vector<int> dof_off, cell_off, slice_off, n_active;
  std::vector<long long> nnz_off, ws_off;
  int total_dof = 0, total_cells = 0, total_slices = 0, n_dir = 0, n_colors = 0;
  long long total_nnz = 0, total_slots = 0;
  std::vector<int> color_ptr;   // cells of colour k: d_cells[color_ptr[k] .. color_ptr[k+1])
  int nq = 0;
  double time = 0;
  int block = 0;
  // device
  int *d_rp = nullptr, *d_col = nullptr;
  double *d_A = nullptr, *d_M = nullptr, *d_S = nullptr;
  int *d_sptr = nullptr, *d_srow = nullptr, *d_scol = nullptr, *d_ssrc = nullptr, *d_slice_prob = nullptr;
  double *d_S64 = nullptr, *d_M64 = nullptr;
  float* d_S32 = nullptr;
  double* d_dinv = nullptr;
  float* d_dinv32 = nullptr;
  int* d_zero = nullptr;
  double* d_geo = nullptr;
  int *d_dofs = nullptr, *d_cells = nullptr, *d_cell_prob = nullptr, *d_active_end = nullptr;
  double* d_coef = nullptr;
  double* d_xyz = nullptr;
  int* d_dir = nullptr;
  uint8_t* d_row_active = nullptr;
  double* d_ref = nullptr;
  int *d_order = nullptr, *d_dof_off = nullptr, *d_n_active = nullptr, *d_slice_off = nullptr;
  long long* d_ws_off = nullptr;
  double *d_x = nullptr, *d_f = nullptr, *d_rhs = nullptr, *d_r = nullptr;
  float *d_w = nullptr, *d_V = nullptr;
  int *d_iters = nullptr, *d_outer = nullptr, *d_fail = nullptr;
  double* d_res = nullptr;
  double* d_time = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // workgroups per problem (BiCGStab), set on the first run
  bool assigned = false;
  int n_wg = 0;
  std::vector<int> groups;
  int *d_wg_prob = nullptr, *d_wg_rank = nullptr, *d_groups = nullptr, *d_bar_timeout = nullptr;
  unsigned int* d_bar = nullptr;
  double* d_part = nullptr;
  // snapshots of all problems (total_dof x snap_max, column-major) and the
  // POD of every problem (basis total_dof x pod_kmax, mean total_dof)
  double* d_snaps = nullptr;
  int snap_max = 0, snap_count = 0;
  double *d_pod_basis = nullptr, *d_pod_mean = nullptr;
  int pod_m = 0, pod_kmax = 0;
  std::vector<int> pod_rank;
  std::vector<double> pod_eigs;   // n_prob x pod_m, largest first

  ~pmx_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : std::initializer_list<void*>{
             d_rp, d_col, d_A, d_M, d_S, d_sptr, d_srow, d_scol, d_ssrc, d_slice_prob, d_S64, d_M64, d_S32,
             d_dinv, d_dinv32, d_zero, d_geo, d_dofs, d_cells, d_cell_prob, d_active_end, d_coef, d_xyz,
             d_dir, d_row_active, d_ref, d_order, d_dof_off, d_n_active, d_slice_off, d_ws_off, d_x, d_f,
             d_rhs, d_r, d_w, d_V, d_iters, d_outer, d_fail, d_res, d_time, d_snaps, d_pod_basis,
             d_pod_mean, d_wg_prob, d_wg_rank, d_groups, d_bar_timeout, d_bar, d_part})
      dfree(p);
  }

  pmx::AssemblyData assembly(int cell_begin, int n_cells) const {
    pmx::AssemblyData ad;
    ad.dim = 2;
    ad.ndof = 3;
    ad.nq = nq;
    ad.n_cells = n_cells;
    ad.ref = d_ref;
    ad.geo = d_geo + (size_t)cell_begin * pmx::kGeoStride;
    ad.dofs = d_dofs + (size_t)cell_begin * 3;
    return ad;
  }

  pmx::CDCoeffs coeffs(bool device_time) const {
    pmx::CDCoeffs co;
    co.example = prm.example;
    co.time = time;
    co.prob_coef = d_coef;
    co.cell_prob = d_cell_prob;
    co.prob_active_end = d_active_end;
    co.time_dev = device_time ? d_time : nullptr;
    return co;
  }

  pmx::DevCSR csr(double* val) const {
    pmx::DevCSR A{};
    A.n = total_dof;
    A.m = total_dof;
    A.nnz = (int)total_nnz;
    A.active = total_dof;
    A.row_ptr = d_rp;
    A.col = d_col;
    A.val = val;
    return A;
  }

  pmx::BatchDev dev() const {
    pmx::BatchDev B;
    B.n_prob = n_prob;
    B.order = d_order;
    B.dof_off = d_dof_off;
    B.n_active = d_n_active;
    B.slice_off = d_slice_off;
    B.ws_off = d_ws_off;
    B.sptr = d_sptr;
    B.srow = d_srow;
    B.scol = d_scol;
    B.S64 = d_S64;
    B.S32 = d_S32;
    B.dinv32 = d_dinv32;
    B.x = d_x;
    B.rhs = d_rhs;
    B.r = d_r;
    B.w = d_w;
    B.V = d_V;
    B.iters = d_iters;
    B.outer = d_outer;
    B.res = d_res;
    B.fail = d_fail;
    if (d_wg_prob) {
      B.n_wg = n_wg;
      B.wg_prob = d_wg_prob;
      B.wg_rank = d_wg_rank;
      B.groups = d_groups;
      B.bar = d_bar;
      B.part = d_part;
      B.bar_timeout = d_bar_timeout;
    }
    return B;
  }

  pmx::BatchSolveParams solve_params() const {
    pmx::BatchSolveParams P;
    P.method = prm.inner_solver;
    P.max_inner = prm.max_inner;
    P.m = prm.gmres_restart;
    P.max_outer = prm.max_outer;
    P.inner_red = prm.inner_reduction;
    P.red = prm.residual_reduction;
    P.tol = prm.residual_tolerance;
    return P;
  }

  // one backward-Euler step of every problem on the context stream
  void step(int blk, hipStream_t st) const {
    pmx::add_scalar(d_time, prm.tau, st);
    pmx::vec_fill(total_dof, 0.0, d_f, st);
    const pmx::CDCoeffs co = coeffs(true);
    const pmx::DevCSR none{};
    for (int k = 0; k < n_colors; ++k) {
      const int n = color_ptr[k + 1] - color_ptr[k];
      pmx::assemble_tcd(assembly(0, total_cells), co, d_cells + color_ptr[k], n, 0, none, none, d_f, st);
    }
    pmx::batch_dirichlet_values(prm.example, d_time, n_dir, d_dir, d_xyz, d_f, st);
    pmx::batch_rhs(total_slices, d_slice_prob, d_dof_off, d_sptr, d_srow, d_scol, d_M64, prm.tau, d_f, d_x,
                   d_rhs, st);
    pmx::batch_dirichlet_copy(n_dir, d_dir, d_f, d_rhs, d_x, st);
    pmx::batch_solve(dev(), solve_params(), blk, st);
  }
};

namespace {

int batch_block(int b) { return (b == 256 || b == 512 || b == 1024) ? b : 0; }

void build_host(pmx_batch* B) {
  const int np = B->n_prob;
  B->meshes.resize(np);
  B->spaces.resize(np);
  B->pats.resize(np);
  pmx::BoundaryConditions bc;
  bc.cond2d = [](int, double) { return (int)pmx::BC_DIRICHLET; };
  bc.cond3d = [](double, double, double) { return (int)pmx::BC_DIRICHLET; };
  std::vector<std::string> err(np);
#pragma omp parallel for schedule(dynamic, 1)
  for (int p = 0; p < np; ++p) {
    try {
      const int n = B->probs[p].n_side;
      int k = 0;
      while ((1 << k) < n) ++k;
      if ((1 << k) == n) B->meshes[p] = pmx::build_mesh(pmx::GEO_UNIT_SQUARE_TRI, k);
      else B->meshes[p] = pmx::build_structured_tri_mesh(n);
      const int level = (int)B->meshes[p].levels.size() - 1;
      B->spaces[p] = pmx::build_space(B->meshes[p], level, 1, bc);
      B->pats[p] = pmx::build_structure(B->spaces[p]);
    } catch (const std::exception& e) {
      err[p] = e.what();
    }
  }
  for (int p = 0; p < np; ++p)
    if (!err[p].empty()) throw PmxError(PMX_E_ARG, "problem " + std::to_string(p) + ": " + err[p]);
  // spaces keep a pointer to their mesh: re-point after the vector is final
  for (int p = 0; p < np; ++p) B->spaces[p].mesh = &B->meshes[p];
}

}  // namespace

extern "C" {

int pmx_batch_default_params(pmx_batch_params* p) {
  return guarded([&] {
    require(p, "NULL argument");
    p->example = PMX_EX_TCD_EXP;
    p->scheme = PMX_TIME_BACKWARD_EULER;
    p->tau = 0.01;
    p->t0 = 0.0;
    p->inner_solver = PMX_BATCH_INNER_BICGSTAB;
    p->gmres_restart = 30;
    p->max_inner = 2000;
    p->max_outer = 50;
    p->inner_reduction = 1e-2;
    p->residual_reduction = 1e-10;
    p->residual_tolerance = 0.0;
    p->block = 0;
    p->max_groups = 16;
  });
}

int pmx_batch_create(pmx_ctx* ctx, int n_problems, const pmx_batch_problem* probs, const pmx_batch_params* prm,
                     pmx_batch** out) {
  return guarded([&] {
    require(ctx && probs && prm && out, "NULL argument");
    require(n_problems > 0, "no problems");
    require(prm->scheme == PMX_TIME_BACKWARD_EULER, "the batched solve implements backward Euler");
    require(prm->example >= 0 && prm->example <= PMX_EX_TCD_LINEAR, "unknown example");
    require(prm->tau > 0, "time step length must be positive");
    require(prm->gmres_restart >= 1 && prm->gmres_restart <= 64, "gmres_restart must be in [1,64]");
    require(prm->max_outer >= 1, "max_outer must be positive");
    require(prm->inner_solver == PMX_BATCH_INNER_GMRES || prm->inner_solver == PMX_BATCH_INNER_BICGSTAB,
            "unknown inner solver");
    require(prm->max_inner >= 1, "max_inner must be positive");
    require(prm->block == 0 || batch_block(prm->block), "block must be 0 (tune), 256, 512 or 1024");
    for (int p = 0; p < n_problems; ++p) {
      require(probs[p].n_side >= 2 && probs[p].n_side <= 2048, "n_side must be in [2,2048]");
      require(probs[p].eps > 0, "eps must be positive");
    }
    auto B = std::make_unique<pmx_batch>();
    B->ctx = ctx;
    B->prm = *prm;
    B->n_prob = n_problems;
    B->probs.assign(probs, probs + n_problems);
    B->time = prm->t0;
    B->block = prm->block;
    build_host(B.get());
    const int np = n_problems;
    B->dof_off.assign(np + 1, 0);
    B->cell_off.assign(np + 1, 0);
    B->slice_off.assign(np + 1, 0);
    B->nnz_off.assign(np + 1, 0);
    B->ws_off.assign(np + 1, 0);
    B->n_active.assign(np, 0);
    const int m1 = (prm->inner_solver == PMX_BATCH_INNER_GMRES) ? prm->gmres_restart + 1 : 7;
    for (int p = 0; p < np; ++p) {
      const pmx::FESpace& fs = B->spaces[p];
      B->n_active[p] = fs.active_bound;
      B->dof_off[p + 1] = B->dof_off[p] + fs.n_dof;
      B->cell_off[p + 1] = B->cell_off[p] + fs.n_cells;
      B->slice_off[p + 1] = B->slice_off[p] + (fs.active_bound + 63) / 64;
      B->nnz_off[p + 1] = B->nnz_off[p] + (long long)B->pats[p].col.size();
      B->ws_off[p + 1] = B->ws_off[p] + (long long)m1 * fs.active_bound;
    }
    B->total_dof = B->dof_off[np];
    B->total_cells = B->cell_off[np];
    B->total_slices = B->slice_off[np];
    B->total_nnz = B->nnz_off[np];
    require(B->total_nnz < (1LL << 31), "batch too large (nnz)");
    // concatenated CSR (global numbering), per-problem SELL slices (local columns)
    std::vector<int> rp(B->total_dof + 1), col(B->total_nnz);
    std::vector<uint8_t> row_active(B->total_dof);
    std::vector<int> sptr(B->total_slices + 1, 0), slice_prob(B->total_slices);
    for (int p = 0; p < np; ++p) {
      const pmx::CSRPattern& P = B->pats[p];
      const int off = B->dof_off[p];
      for (int i = 0; i < P.n_rows; ++i) {
        rp[off + i] = (int)(B->nnz_off[p] + P.row_ptr[i]);
        row_active[off + i] = i < P.active_bound;
      }
      for (size_t k = 0; k < P.col.size(); ++k) col[B->nnz_off[p] + k] = off + P.col[k];
      for (int s = B->slice_off[p]; s < B->slice_off[p + 1]; ++s) {
        slice_prob[s] = p;
        const int r0 = (s - B->slice_off[p]) * 64;
        int width = 0;
        for (int l = 0; l < 64 && r0 + l < P.active_bound; ++l)
          width = std::max(width, P.row_ptr[r0 + l + 1] - P.row_ptr[r0 + l]);
        sptr[s + 1] = width * 64;
      }
    }
    rp[B->total_dof] = (int)B->total_nnz;
    for (int s = 0; s < B->total_slices; ++s) sptr[s + 1] += sptr[s];
    B->total_slots = sptr[B->total_slices];
    std::vector<int> srow((size_t)B->total_slices * 64, -1), scol(B->total_slots), ssrc(B->total_slots, -1);
#pragma omp parallel for schedule(dynamic, 4)
    for (int p = 0; p < np; ++p) {
      const pmx::CSRPattern& P = B->pats[p];
      for (int s = B->slice_off[p]; s < B->slice_off[p + 1]; ++s) {
        const int r0 = (s - B->slice_off[p]) * 64;
        const int width = (sptr[s + 1] - sptr[s]) / 64;
        for (int l = 0; l < 64; ++l) {
          const int r = r0 + l;
          const bool valid = r < P.active_bound;
          if (valid) srow[(size_t)s * 64 + l] = r;
          for (int j = 0; j < width; ++j) {
            const size_t k = (size_t)sptr[s] + (size_t)j * 64 + l;
            if (valid && j < P.row_ptr[r + 1] - P.row_ptr[r]) {
              scol[k] = P.col[P.row_ptr[r] + j];
              ssrc[k] = (int)(B->nnz_off[p] + P.row_ptr[r] + j);
            } else {
              scol[k] = valid ? r : 0;
            }
          }
        }
      }
    }
    // geometry, DOF maps, colours, coefficients, Dirichlet list
    std::vector<double> geo((size_t)B->total_cells * pmx::kGeoStride, 0.0);
    std::vector<int> dofs((size_t)B->total_cells * 3), cell_prob(B->total_cells), active_end(np);
    std::vector<double> coef((size_t)np * 4), xyz((size_t)B->total_dof * 3);
    std::vector<int> dir;
    std::vector<std::vector<int>> by_color;
    for (int p = 0; p < np; ++p) {
      const pmx::FESpace& fs = B->spaces[p];
      const pmx::Mesh& m = B->meshes[p];
      const int c0 = B->cell_off[p], off = B->dof_off[p];
      active_end[p] = off + fs.active_bound;
      const double* bb = B->probs[p].b;
      coef[4 * p] = B->probs[p].eps;
      coef[4 * p + 1] = bb[0];
      coef[4 * p + 2] = bb[1];
      coef[4 * p + 3] = B->probs[p].c;
      for (int c = 0; c < fs.n_cells; ++c) {
        const pmx::AffineMap a = pmx::affine_map(m, fs.level, c);
        double* g = &geo[(size_t)(c0 + c) * pmx::kGeoStride];
        g[0] = a.x0[0];
        g[1] = a.x0[1];
        for (int k = 0; k < 4; ++k) { g[3 + k] = a.J[k]; g[12 + k] = a.Jinv[k]; }
        g[21] = a.detJ;
        g[22] = pmx::cell_diameter(m, fs.level, c);
        for (int k = 0; k < 3; ++k) dofs[(size_t)(c0 + c) * 3 + k] = off + fs.global_numbers[(size_t)c * 3 + k];
        cell_prob[c0 + c] = p;
      }
      std::memcpy(&xyz[(size_t)off * 3], fs.dof_xyz.data(), (size_t)fs.n_dof * 3 * sizeof(double));
      for (int i = fs.active_bound; i < fs.n_dof; ++i) dir.push_back(off + i);
      std::vector<int> order, cptr;
      pmx::color_cells(fs, order, cptr);
      if ((int)by_color.size() < (int)cptr.size() - 1) by_color.resize(cptr.size() - 1);
      for (size_t k = 0; k + 1 < cptr.size(); ++k)
        for (int i = cptr[k]; i < cptr[k + 1]; ++i) by_color[k].push_back(c0 + order[i]);
    }
    B->n_dir = (int)dir.size();
    B->n_colors = (int)by_color.size();
    std::vector<int> cells;
    B->color_ptr.assign(1, 0);
    for (auto& v : by_color) {
      cells.insert(cells.end(), v.begin(), v.end());
      B->color_ptr.push_back((int)cells.size());
    }
    // workgroup order: largest problem first
    std::vector<int> order(np);
    for (int p = 0; p < np; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return B->n_active[a] > B->n_active[b]; });
    // reference table of the P1 triangle (values, gradients, second derivatives)
    const pmx::Element el = pmx::make_element(pmx::SHAPE_TRI, 1);
    const pmx::QuadRule rule = pmx::assembly_rule(el, 0);
    B->nq = rule.n;
    const int rs = 1 + 2 + 3 * (1 + 2 + 3);
    std::vector<double> ref((size_t)rs * rule.n, 0.0);
    for (int q = 0; q < rule.n; ++q) {
      double* r = &ref[(size_t)q * rs];
      r[0] = rule.w[q];
      r[1] = rule.xi[2 * q];
      r[2] = rule.xi[2 * q + 1];
      el.eval(&rule.xi[2 * q], r + 3, r + 3 + 3, r + 3 + 3 * 3);
    }

    hipStream_t st = ctx->stream;
    B->d_rp = upload(rp.data(), rp.size(), st);
    B->d_col = upload(col.data(), col.size(), st);
    B->d_A = dmalloc<double>(B->total_nnz);
    B->d_M = dmalloc<double>(B->total_nnz);
    B->d_S = dmalloc<double>(B->total_nnz);
    B->d_sptr = upload(sptr.data(), sptr.size(), st);
    B->d_srow = upload(srow.data(), srow.size(), st);
    B->d_scol = upload(scol.data(), scol.size(), st);
    B->d_ssrc = upload(ssrc.data(), ssrc.size(), st);
    B->d_slice_prob = upload(slice_prob.data(), slice_prob.size(), st);
    B->d_S64 = dmalloc<double>(B->total_slots);
    B->d_M64 = dmalloc<double>(B->total_slots);
    B->d_S32 = dmalloc<float>(B->total_slots);
    B->d_dinv = dmalloc<double>(B->total_dof);
    B->d_dinv32 = dmalloc<float>(B->total_dof);
    B->d_zero = dmalloc<int>(1);
    B->d_geo = upload(geo.data(), geo.size(), st);
    B->d_dofs = upload(dofs.data(), dofs.size(), st);
    B->d_cells = upload(cells.data(), cells.size(), st);
    B->d_cell_prob = upload(cell_prob.data(), cell_prob.size(), st);
    B->d_active_end = upload(active_end.data(), active_end.size(), st);
    B->d_coef = upload(coef.data(), coef.size(), st);
    B->d_xyz = upload(xyz.data(), xyz.size(), st);
    B->d_dir = dir.empty() ? nullptr : upload(dir.data(), dir.size(), st);
    B->d_row_active = upload(row_active.data(), row_active.size(), st);
    B->d_ref = upload(ref.data(), ref.size(), st);
    B->d_order = upload(order.data(), order.size(), st);
    B->d_dof_off = upload(B->dof_off.data(), B->dof_off.size(), st);
    B->d_n_active = upload(B->n_active.data(), B->n_active.size(), st);
    B->d_slice_off = upload(B->slice_off.data(), B->slice_off.size(), st);
    B->d_ws_off = upload(B->ws_off.data(), B->ws_off.size(), st);
    B->d_x = dmalloc<double>(B->total_dof);
    B->d_f = dmalloc<double>(B->total_dof);
    B->d_rhs = dmalloc<double>(B->total_dof);
    B->d_r = dmalloc<double>(B->total_dof);
    B->d_w = dmalloc<float>(B->total_dof);
    B->d_V = dmalloc<float>((size_t)std::max<long long>(B->ws_off[np], 1));
    B->d_iters = dmalloc<int>(np);
    B->d_outer = dmalloc<int>(np);
    B->d_fail = dmalloc<int>(np);
    B->d_res = dmalloc<double>(np);
    B->d_time = dmalloc<double>(1);
    HIP_CHECK(hipStreamSynchronize(st));
    *out = B.release();
  });
}

int pmx_batch_destroy(pmx_batch* b) {
  return guarded([&] { delete b; });
}

int pmx_batch_get_info(const pmx_batch* b, pmx_batch_info* out) {
  return guarded([&] {
    require(b && out, "NULL argument");
    out->n_problems = b->n_prob;
    out->total_dof = b->total_dof;
    out->total_active = 0;
    for (int p = 0; p < b->n_prob; ++p) out->total_active += b->n_active[p];
    out->total_cells = b->total_cells;
    out->total_nnz = b->total_nnz;
    out->block = b->block;
    out->time = b->time;
    out->n_workgroups = b->d_wg_prob ? b->n_wg : b->n_prob;
  });
}

int pmx_batch_problem_size(const pmx_batch* b, int problem, int* n_dof, int* active_bound, int* nnz) {
  return guarded([&] {
    require(b && problem >= 0 && problem < b->n_prob, "invalid problem");
    if (n_dof) *n_dof = b->spaces[problem].n_dof;
    if (active_bound) *active_bound = b->n_active[problem];
    if (nnz) *nnz = (int)b->pats[problem].col.size();
  });
}

int pmx_batch_initialize(pmx_batch* b) {
  return guarded([&] {
    require(b, "NULL argument");
    hipStream_t st = b->ctx->stream;
    if (b->exec) { HIP_CHECK(hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
    if (b->graph) { HIP_CHECK(hipGraphDestroy(b->graph)); b->graph = nullptr; }
    b->time = b->prm.t0;
    HIP_CHECK(hipMemcpyAsync(b->d_time, &b->time, sizeof(double), hipMemcpyHostToDevice, st));
    // u = interpolant of the initial condition
    pmx::interpolate_example(b->prm.example, b->time, b->total_dof, b->d_xyz, b->d_x, st);
    // stiffness and mass matrices (time-independent coefficients)
    pmx::vec_fill((int)b->total_nnz, 0.0, b->d_A, st);
    pmx::vec_fill((int)b->total_nnz, 0.0, b->d_M, st);
    const pmx::CDCoeffs co = b->coeffs(false);
    const pmx::DevCSR A = b->csr(b->d_A), M = b->csr(b->d_M);
    for (int k = 0; k < b->n_colors; ++k) {
      const int n = b->color_ptr[k + 1] - b->color_ptr[k];
      pmx::assemble_tcd(b->assembly(0, b->total_cells), co, b->d_cells + b->color_ptr[k], n, 0, A, M, nullptr,
                        st);
    }
    // S = tau A + M on the active rows, 1/diag(S), slot copies
    pmx::batch_system(b->total_dof, b->d_rp, b->d_row_active, b->prm.tau, b->d_A, b->d_M, b->d_S, st);
    HIP_CHECK(hipMemsetAsync(b->d_zero, 0, sizeof(int), st));
    pmx::csr_inv_diag(b->csr(b->d_S), b->d_dinv, b->d_zero, st);
    pmx::vec_f64_to_f32(b->total_dof, b->d_dinv, b->d_dinv32, st);
    pmx::sell_gather(b->total_slots, b->d_ssrc, b->d_S, b->d_S64, b->d_S32, st);
    pmx::sell_gather(b->total_slots, b->d_ssrc, b->d_M, b->d_M64, nullptr, st);
    HIP_CHECK(hipMemsetAsync(b->d_iters, 0, sizeof(int) * b->n_prob, st));
    HIP_CHECK(hipMemsetAsync(b->d_outer, 0, sizeof(int) * b->n_prob, st));
    HIP_CHECK(hipMemsetAsync(b->d_fail, 0, sizeof(int) * b->n_prob, st));
    HIP_CHECK(hipMemsetAsync(b->d_res, 0, sizeof(double) * b->n_prob, st));
    HIP_CHECK(hipGetLastError());
    int zero = 0;
    HIP_CHECK(hipMemcpyAsync(&zero, b->d_zero, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (zero) throw PmxError(PMX_E_ZERO_DIAG, "zero diagonal entry in a batched system matrix");
  });
}

// Workgroups per problem for the inner solver's kernel from the measured
// iterations of one step: cost = iterations x slices; one more workgroup goes
// to the problem of the largest cost per workgroup while that exceeds the
// mean cost per workgroup of the grid, within the number of workgroups
// that can be resident at once (the problem's workgroups wait for each other)
// and max_groups per problem.
static void clear_groups(pmx_batch* b) {
  dfree(b->d_wg_prob); dfree(b->d_wg_rank); dfree(b->d_groups); dfree(b->d_bar); dfree(b->d_part);
  dfree(b->d_bar_timeout);
  b->groups.clear();
  b->n_wg = 0;
}

static void assign_groups(pmx_batch* b, const std::vector<int>& iters) {
  clear_groups(b);
  const int np = b->n_prob;
  const int gmax = std::min(std::max(1, b->prm.max_groups), pmx::kBatchMaxGroups);
  if (gmax <= 1) return;
  const int capacity =
      pmx::batch_solve_capacity(b->block, b->prm.inner_solver == PMX_BATCH_INNER_BICGSTAB ? 1 : 0);
  if (capacity <= np) return;
  std::vector<double> cost(np);
  double total = 0;
  for (int q = 0; q < np; ++q) {
    cost[q] = (double)std::max(1, iters[q]) * (double)(b->slice_off[q + 1] - b->slice_off[q]);
    total += cost[q];
  }
  std::vector<int> G(np, 1);
  int n_wg = np;
  for (;;) {
    int top = 0;
    for (int q = 1; q < np; ++q)
      if (cost[q] / G[q] > cost[top] / G[top]) top = q;
    if (n_wg >= capacity || G[top] >= gmax || cost[top] / G[top] <= 1.0 * total / n_wg) break;
    ++G[top];
    ++n_wg;
  }
  if (n_wg == np) return;
  // workgroups of the costliest problems first, ranks contiguous
  std::vector<int> order(np);
  for (int q = 0; q < np; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  std::vector<int> wg_prob, wg_rank;
  for (int q : order)
    for (int g = 0; g < G[q]; ++g) {
      wg_prob.push_back(q);
      wg_rank.push_back(g);
    }
  hipStream_t st = b->ctx->stream;
  b->groups = G;
  b->n_wg = n_wg;
  b->d_wg_prob = upload(wg_prob.data(), wg_prob.size(), st);
  b->d_wg_rank = upload(wg_rank.data(), wg_rank.size(), st);
  b->d_groups = upload(G.data(), G.size(), st);
  b->d_bar = dmalloc<unsigned int>(np);
  b->d_part = dmalloc<double>((size_t)np * 2 * pmx::kBatchMaxGroups * pmx::kBatchPartStride);
  b->d_bar_timeout = dmalloc<int>(1);
  HIP_CHECK(hipMemsetAsync(b->d_bar_timeout, 0, sizeof(int), st));
  HIP_CHECK(hipStreamSynchronize(st));
}

int pmx_batch_run(pmx_batch* b, int n_steps, pmx_batch_report* rep) {
  return guarded([&] {
    require(b && n_steps >= 0, "invalid argument");
    hipStream_t st = b->ctx->stream;
    // the state every trial step starts from and is restored to
    auto save = [&](double* x0, std::vector<int>& it0) {
      HIP_CHECK(hipMemcpyAsync(x0, b->d_x, sizeof(double) * b->total_dof, hipMemcpyDeviceToDevice, st));
      HIP_CHECK(hipMemcpyAsync(it0.data(), b->d_iters, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
    };
    auto restore = [&](const double* x0, const std::vector<int>& it0) {
      HIP_CHECK(hipMemcpyAsync(b->d_time, &b->time, sizeof(double), hipMemcpyHostToDevice, st));
      HIP_CHECK(hipMemcpyAsync(b->d_x, x0, sizeof(double) * b->total_dof, hipMemcpyDeviceToDevice, st));
      HIP_CHECK(hipMemcpyAsync(b->d_iters, it0.data(), sizeof(int) * b->n_prob, hipMemcpyHostToDevice, st));
      HIP_CHECK(hipMemsetAsync(b->d_outer, 0, sizeof(int) * b->n_prob, st));
      HIP_CHECK(hipMemsetAsync(b->d_fail, 0, sizeof(int) * b->n_prob, st));
      HIP_CHECK(hipStreamSynchronize(st));
    };
    if (!b->assigned) {
      // one step with one workgroup per problem measures every problem's
      // iterations; without a given workgroup size the candidates (size,
      // with and without workgroups shared by the costliest problems) are
      // timed on one step each from the same state and the fastest is kept
      const int given = b->block;
      double* x0 = dmalloc<double>(b->total_dof);
      std::vector<int> it0(b->n_prob), it1(b->n_prob);
      save(x0, it0);
      clear_groups(b);
      b->block = given ? given : 1024;
      HIP_CHECK(hipMemcpyAsync(b->d_time, &b->time, sizeof(double), hipMemcpyHostToDevice, st));
      b->step(b->block, st);
      HIP_CHECK(hipMemcpyAsync(it1.data(), b->d_iters, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      for (int q = 0; q < b->n_prob; ++q) it1[q] -= it0[q];
      restore(x0, it0);
      hipEvent_t e0, e1;
      HIP_CHECK(hipEventCreate(&e0));
      HIP_CHECK(hipEventCreate(&e1));
      float best = 1e30f;
      int best_blk = b->block;
      bool best_grouped = true;
      // a given size is not timed (the choice stays deterministic): the
      // costliest problems share workgroups whenever the grid fits
      std::vector<int> blocks = given ? std::vector<int>{} : std::vector<int>{256, 512, 1024};
      for (int blk : blocks)
        for (int grouped = 0; grouped < 2; ++grouped) {
          b->block = blk;
          if (grouped) {
            assign_groups(b, it1);
            if (!b->d_wg_prob) continue;   // nothing to share at this size
          } else {
            clear_groups(b);
          }
          HIP_CHECK(hipEventRecord(e0, st));
          b->step(blk, st);
          HIP_CHECK(hipEventRecord(e1, st));
          HIP_CHECK(hipEventSynchronize(e1));
          float ms = 0;
          HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          int flag = 0;
          if (b->d_bar_timeout) HIP_CHECK(hipMemcpy(&flag, b->d_bar_timeout, sizeof(int), hipMemcpyDeviceToHost));
          if (!flag && ms < best) { best = ms; best_blk = blk; best_grouped = grouped != 0; }
          restore(x0, it0);
        }
      b->block = best_blk;
      if (best_grouped) assign_groups(b, it1);
      else clear_groups(b);
      b->assigned = true;
      HIP_CHECK(hipEventDestroy(e0));
      HIP_CHECK(hipEventDestroy(e1));
      dfree(x0);
    }
    auto capture = [&]() {
      if (b->exec) { HIP_CHECK(hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
      if (b->graph) { HIP_CHECK(hipGraphDestroy(b->graph)); b->graph = nullptr; }
      HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      b->step(b->block, st);
      HIP_CHECK(hipStreamEndCapture(st, &b->graph));
      HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    };
    if (!b->exec) capture();
    // Shared workgroups wait for each other at barriers, which assumes that
    // all workgroups of the grid are resident at once (occupancy x CUs of an
    // otherwise idle device).  When a barrier gives up (e.g. another process
    // holds CUs), the replayed steps are redone from the saved state with one
    // workgroup per problem, which needs no barrier.  The flag is cleared
    // before every replay, outside graph capture.
    double* x_save = nullptr;
    std::vector<int> it_save(b->n_prob), outer_save(b->n_prob), fail_save(b->n_prob);
    if (b->d_bar_timeout) {
      HIP_CHECK(hipMemsetAsync(b->d_bar_timeout, 0, sizeof(int), st));
      x_save = dmalloc<double>(b->total_dof);
      HIP_CHECK(hipMemcpyAsync(x_save, b->d_x, sizeof(double) * b->total_dof, hipMemcpyDeviceToDevice, st));
      HIP_CHECK(hipMemcpyAsync(it_save.data(), b->d_iters, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipMemcpyAsync(outer_save.data(), b->d_outer, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipMemcpyAsync(fail_save.data(), b->d_fail, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
    }
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < n_steps; ++k) HIP_CHECK(hipGraphLaunch(b->exec, st));
    HIP_CHECK(hipStreamSynchronize(st));
    auto t1 = std::chrono::steady_clock::now();
    HIP_CHECK(hipGetLastError());
    if (b->d_bar_timeout) {
      int flag = 0;
      HIP_CHECK(hipMemcpy(&flag, b->d_bar_timeout, sizeof(int), hipMemcpyDeviceToHost));
      if (flag) {
        HIP_CHECK(hipMemcpyAsync(b->d_time, &b->time, sizeof(double), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->d_x, x_save, sizeof(double) * b->total_dof, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->d_iters, it_save.data(), sizeof(int) * b->n_prob, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->d_outer, outer_save.data(), sizeof(int) * b->n_prob, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(b->d_fail, fail_save.data(), sizeof(int) * b->n_prob, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipStreamSynchronize(st));
        clear_groups(b);
        capture();
        t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < n_steps; ++k) HIP_CHECK(hipGraphLaunch(b->exec, st));
        HIP_CHECK(hipStreamSynchronize(st));
        t1 = std::chrono::steady_clock::now();
        HIP_CHECK(hipGetLastError());
      }
      dfree(x_save);
    }
    b->time += n_steps * b->prm.tau;
    if (rep) {
      std::vector<int> it(b->n_prob), fail(b->n_prob), outer(b->n_prob);
      HIP_CHECK(hipMemcpy(it.data(), b->d_iters, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost));
      HIP_CHECK(hipMemcpy(outer.data(), b->d_outer, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost));
      HIP_CHECK(hipMemcpy(fail.data(), b->d_fail, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost));
      rep->seconds = std::chrono::duration<double>(t1 - t0).count();
      rep->total_inner_iterations = 0;
      rep->total_outer_iterations = 0;
      rep->max_inner_iterations = 0;
      rep->failed_solves = 0;
      for (int p = 0; p < b->n_prob; ++p) {
        rep->total_inner_iterations += it[p];
        rep->total_outer_iterations += outer[p];
        rep->max_inner_iterations = std::max(rep->max_inner_iterations, it[p]);
        rep->failed_solves += fail[p];
      }
      rep->block = b->block;
    }
  });
}

int pmx_batch_get_solution(const pmx_batch* b, int problem, double* host_u) {
  return guarded([&] {
    require(b && host_u && problem >= 0 && problem < b->n_prob, "invalid argument");
    const int n = b->dof_off[problem + 1] - b->dof_off[problem];
    HIP_CHECK(hipMemcpyAsync(host_u, b->d_x + b->dof_off[problem], sizeof(double) * n, hipMemcpyDeviceToHost,
                             b->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  });
}

int pmx_batch_get_iterations(const pmx_batch* b, int* host_inner, int* host_outer, int* host_failed) {
  return guarded([&] {
    require(b, "NULL argument");
    hipStream_t st = b->ctx->stream;
    if (host_inner) HIP_CHECK(hipMemcpyAsync(host_inner, b->d_iters, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
    if (host_outer) HIP_CHECK(hipMemcpyAsync(host_outer, b->d_outer, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
    if (host_failed) HIP_CHECK(hipMemcpyAsync(host_failed, b->d_fail, sizeof(int) * b->n_prob, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pmx_batch_get_structure(const pmx_batch* b, int problem, int* host_row_ptr, int* host_col,
                            int* host_global_numbers) {
  return guarded([&] {
    require(b && problem >= 0 && problem < b->n_prob, "invalid argument");
    const pmx::CSRPattern& P = b->pats[problem];
    if (host_row_ptr) std::memcpy(host_row_ptr, P.row_ptr.data(), P.row_ptr.size() * sizeof(int));
    if (host_col) std::memcpy(host_col, P.col.data(), P.col.size() * sizeof(int));
    if (host_global_numbers)
      std::memcpy(host_global_numbers, b->spaces[problem].global_numbers.data(),
                  b->spaces[problem].global_numbers.size() * sizeof(int));
  });
}

int pmx_batch_get_matrices(const pmx_batch* b, int problem, double* host_stiffness, double* host_mass,
                           double* host_system) {
  return guarded([&] {
    require(b && problem >= 0 && problem < b->n_prob, "invalid argument");
    const size_t n = b->pats[problem].col.size();
    const size_t o = b->nnz_off[problem];
    hipStream_t st = b->ctx->stream;
    if (host_stiffness) HIP_CHECK(hipMemcpyAsync(host_stiffness, b->d_A + o, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (host_mass) HIP_CHECK(hipMemcpyAsync(host_mass, b->d_M + o, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (host_system) HIP_CHECK(hipMemcpyAsync(host_system, b->d_S + o, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  });
}

int pmx_batch_errors(pmx_batch* b, int problem, double* host_errors4) {
  return guarded([&] {
    require(b && host_errors4 && problem >= 0 && problem < b->n_prob, "invalid argument");
    hipStream_t st = b->ctx->stream;
    const int c0 = b->cell_off[problem], nc = b->cell_off[problem + 1] - c0;
    const pmx::AssemblyData ad = b->assembly(c0, nc);
    pmx::CDCoeffs co;
    co.example = b->prm.example;
    co.time = b->time;
    co.eps = b->probs[problem].eps;
    co.b[0] = b->probs[problem].b[0];
    co.b[1] = b->probs[problem].b[1];
    co.c = b->probs[problem].c;
    const int nb = (nc + 255) / 256;
    double* partial = dmalloc<double>((size_t)nb * 4);
    pmx::cd_errors(ad, co, b->d_x, partial, b->ctx->scal, st);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(host_errors4, b->ctx->scal, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    dfree(partial);
  });
}

// ------------------------------------------------ snapshots and POD per problem
int pmx_batch_snapshots_reserve(pmx_batch* b, int max_snaps) {
  return guarded([&] {
    require(b && max_snaps >= 1, "invalid argument");
    dfree(b->d_snaps);
    b->d_snaps = dmalloc<double>((size_t)b->total_dof * max_snaps);
    b->snap_max = max_snaps;
    b->snap_count = 0;
  });
}

// SnapshotsCollector::write_data for every problem: one device copy of the
// current solutions into the next snapshot column
int pmx_batch_write_snapshot(pmx_batch* b) {
  return guarded([&] {
    require(b, "NULL argument");
    if (!b->d_snaps) throw PmxError(PMX_E_STATE, "reserve snapshot storage first");
    if (b->snap_count >= b->snap_max) throw PmxError(PMX_E_ARG, "snapshot storage is full");
    HIP_CHECK(hipMemcpyAsync(b->d_snaps + (size_t)b->snap_count * b->total_dof, b->d_x,
                             sizeof(double) * b->total_dof, hipMemcpyDeviceToDevice, b->ctx->stream));
    b->snap_count++;
  });
}

// POD::compute_basis for every problem: mean and fluctuations of all problems
// in one launch, the Gram matrices of all problems from one launch pair (the
// L2 variant with the block-diagonal mass matrix of the batch), the m x m
// eigenproblems on the host, the bases of all problems in one launch
int pmx_batch_compute_pod(pmx_batch* b, const pmx_pod_params* p) {
  return guarded([&] {
    require(b && p, "NULL argument");
    require(p->pod_inner_product == PMX_POD_EUCLIDEAN || p->pod_inner_product == PMX_POD_L2,
            "pod_inner_product must be euclidean or L2");
    if (b->snap_count < 1) throw PmxError(PMX_E_STATE, "Snapshots are not available");
    hipStream_t st = b->ctx->stream;
    const int n = b->total_dof, m = b->snap_count, np = b->n_prob;
    double* S = dmalloc<double>((size_t)n * m);
    HIP_CHECK(hipMemcpyAsync(S, b->d_snaps, sizeof(double) * n * m, hipMemcpyDeviceToDevice, st));
    dfree(b->d_pod_mean);
    dfree(b->d_pod_basis);
    b->d_pod_mean = dmalloc<double>(n);
    if (p->pod_fluctuations_only) pmx::pod_mean(n, m, n, S, b->d_pod_mean, true, st);
    else pmx::vec_fill(n, 0.0, b->d_pod_mean, st);
    double* WS = nullptr;
    if (p->pod_inner_product == PMX_POD_L2) {
      WS = dmalloc<double>((size_t)n * m);
      pmx::pod_spmm(n, m, n, b->d_rp, b->d_col, b->d_M, S, WS, st);
    }
    std::vector<int> row0(b->dof_off.begin(), b->dof_off.begin() + np), rows(np);
    for (int q = 0; q < np; ++q) rows[q] = b->dof_off[q + 1] - b->dof_off[q];
    std::vector<double> G;
    pod_gram_jobs(st, row0, rows, m, n, S, WS ? WS : S, G);
    b->pod_m = m;
    b->pod_rank.assign(np, 0);
    b->pod_eigs.assign((size_t)np * m, 0.0);
    std::vector<PodEig> eig(np);
#pragma omp parallel for schedule(dynamic)
    for (int q = 0; q < np; ++q) eig[q] = pod_eigen(G.data() + (size_t)q * m * m, m, p->pod_rank);
    int kmax = 1;
    for (int q = 0; q < np; ++q) {
      b->pod_rank[q] = eig[q].rank;
      std::copy(eig[q].eigs.begin(), eig[q].eigs.end(), b->pod_eigs.begin() + (size_t)q * m);
      kmax = std::max(kmax, eig[q].rank);
    }
    b->pod_kmax = kmax;
    std::vector<double> C((size_t)np * m * kmax, 0.0);
    for (int q = 0; q < np; ++q)
      for (int k = 0; k < eig[q].rank; ++k)
        for (int j = 0; j < m; ++j) C[(size_t)q * m * kmax + (size_t)k * m + j] = eig[q].C[(size_t)k * m + j];
    b->d_pod_basis = dmalloc<double>((size_t)n * kmax);
    double* d_C = upload(C.data(), C.size(), st);
    int* d_rank = upload(b->pod_rank.data(), b->pod_rank.size(), st);
    pmx::pod_basis(n, np, b->d_dof_off, d_rank, m, kmax, n, S, d_C, b->d_pod_basis, st);
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipGetLastError());
    dfree(d_C); dfree(d_rank); dfree(S); dfree(WS);
  });
}

int pmx_batch_get_pod(const pmx_batch* b, int problem, int* rank, double* host_eigs, double* host_basis,
                      double* host_mean) {
  return guarded([&] {
    require(b && problem >= 0 && problem < b->n_prob, "invalid argument");
    if (!b->d_pod_mean) throw PmxError(PMX_E_STATE, "POD not computed");
    hipStream_t st = b->ctx->stream;
    const int o = b->dof_off[problem], n = b->dof_off[problem + 1] - o, r = b->pod_rank[problem];
    if (rank) *rank = r;
    if (host_eigs) std::copy(b->pod_eigs.begin() + (size_t)problem * b->pod_m,
                             b->pod_eigs.begin() + (size_t)(problem + 1) * b->pod_m, host_eigs);
    if (host_basis && r > 0)
      HIP_CHECK(hipMemcpy2DAsync(host_basis, sizeof(double) * n, b->d_pod_basis + o, sizeof(double) * b->total_dof,
                                 sizeof(double) * n, r, hipMemcpyDeviceToHost, st));
    if (host_mean)
      HIP_CHECK(hipMemcpyAsync(host_mean, b->d_pod_mean + o, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  });
}

}  // extern "C"
