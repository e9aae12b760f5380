// C ABI of the solve path (include/pmx.h): host-side set-up objects, device
// data and the host-driven Krylov / multigrid loops that launch the HIP
// kernels.  Every numerical operation on vectors and matrices runs on the
// GPU; the host only keeps scalars (Hessenberg matrix, Givens rotations,
// stopping decisions), as the reference keeps them in its solver objects.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hip/examples.hpp"
#include "internal.hpp"

using namespace pmx_internal;
thread_local std::string pmx_internal::g_last_error = "";

namespace pmx_internal {

constexpr int kSellSigma = 4096;

pmx::DevSELL upload_sell(const pmx::SellPattern& sp, hipStream_t st) {
  pmx::DevSELL S;
  S.n_slices = sp.n_slices;
  S.total = sp.src.size();   // slots of either layout (col is empty in the 16-bit one)
  if ((size_t)sp.slice_ptr.back() != S.total || (sp.off.empty() ? sp.col.size() : sp.off.size()) != S.total)
    throw PmxError(PMX_E_ARG, "inconsistent SELL layout");
  S.slice_ptr = upload(sp.slice_ptr.data(), sp.slice_ptr.size(), st);
  S.row = upload(sp.row.data(), sp.row.size(), st);
  S.col = sp.col.empty() ? nullptr : upload(sp.col.data(), sp.col.size(), st);
  S.off = sp.off.empty() ? nullptr : upload(sp.off.data(), sp.off.size(), st);
  S.src = upload(sp.src.data(), sp.src.size(), st);
  S.n_bd = (int)sp.bd_src.size();
  S.bd_ptr = upload(sp.bd_ptr.data(), sp.bd_ptr.size(), st);
  S.bd_src = sp.bd_src.empty() ? nullptr : upload(sp.bd_src.data(), sp.bd_src.size(), st);
  S.bd_col = sp.bd_col.empty() ? nullptr : upload(sp.bd_col.data(), sp.bd_col.size(), st);
  S.pack = sp.pack;
  if (S.pack != 1 && (S.pack != 4 || !S.off)) throw PmxError(PMX_E_ARG, "packed SELL slots need the 16-bit layout");
  if (!sp.slice_pat.empty()) {
    if ((int)sp.slice_pat.size() != sp.n_slices || !S.off)
      throw PmxError(PMX_E_ARG, "inconsistent SELL pattern layout");
    S.slice_pat = upload(sp.slice_pat.data(), sp.slice_pat.size(), st);
    S.pat_off = upload(sp.pat_off.data(), sp.pat_off.size(), st);
    S.n_patterns = sp.n_patterns;
  }
  return S;
}

void attach_sell(pmx_matrix* M, hipStream_t st) {
  const pmx::SellPattern sp = pmx::build_sell(*M->pattern, kSellSigma);
  M->A.sell = upload_sell(sp, st);
  M->A.sell.val = dmalloc<double>(M->A.sell.total);
  M->A.lpr = 0;
  M->sell_valid = false;
}

void ensure_sell_values(pmx_matrix* M, hipStream_t st) {
  if (M->sell_valid || M->A.sell.total == 0) return;
  pmx::sell_fill_f64(M->A, M->A.sell.val, st);
  HIP_CHECK(hipGetLastError());
  M->sell_valid = true;
}

}  // namespace pmx_internal

namespace {

int auto_lanes(int nnz, int n) {
  const double avg = n ? (double)nnz / n : 1.0;
  if (avg > 40) return 64;
  if (avg > 20) return 32;
  if (avg > 10) return 16;
  if (avg > 5) return 8;
  return 4;
}

int builtin_cond2d(int builtin, int comp, double) {
  if (builtin == PMX_BC_SINE_LAPLACE) return comp == 1 ? pmx::BC_NEUMANN : pmx::BC_DIRICHLET;
  return pmx::BC_DIRICHLET;
}

const RefTable& ref_table(pmx_space* s, int hexa_rule) {
  const int key = (s->fs.el.shape == pmx::SHAPE_HEXA) ? hexa_rule : 0;
  auto it = s->ref.find(key);
  if (it != s->ref.end()) return it->second;
  RefTable t;
  const pmx::Element& e = s->fs.el;
  t.rule = pmx::assembly_rule(e, key);
  t.nq = t.rule.n;
  const int dim = e.dim, nd = e.ndof, nc = (dim == 3) ? 6 : 3;
  const int rs = 1 + dim + nd * (1 + dim + nc);
  t.h_ref.assign((size_t)rs * t.nq, 0.0);
  for (int q = 0; q < t.nq; ++q) {
    double* r = &t.h_ref[(size_t)q * rs];
    r[0] = t.rule.w[q];
    for (int d = 0; d < dim; ++d) r[1 + d] = t.rule.xi[q * dim + d];
    e.eval(&t.rule.xi[q * dim], r + 1 + dim, r + 1 + dim + nd, r + 1 + dim + nd * (1 + dim));
  }
  t.d_ref = upload(t.h_ref.data(), t.h_ref.size(), s->ctx->stream);
  return s->ref.emplace(key, std::move(t)).first->second;
}

pmx::AssemblyData assembly_data(pmx_space* s, const RefTable& t) {
  pmx::AssemblyData ad;
  ad.dim = s->fs.el.dim;
  ad.ndof = s->fs.ndof_cell;
  ad.nq = t.nq;
  ad.n_cells = s->fs.n_cells;
  ad.ref = t.d_ref;
  ad.geo = s->d_geo;
  ad.dofs = s->d_dofs;
  return ad;
}

pmx::CDCoeffs cd_coeffs(const pmx_cd_params* p) {
  pmx::CDCoeffs co;
  co.example = p->example;
  co.supg = (p->disc == PMX_DISC_SUPG) ? 1 : (p->disc == PMX_DISC_GLS) ? 2 : 0;
  co.eps = p->eps;
  for (int d = 0; d < 3; ++d) co.b[d] = p->b[d];
  co.c = p->c;
  co.f0 = p->f0;
  co.delta0 = p->delta0;
  co.poly_degree = p->poly_degree;
  co.table = p->dev_coeff_table;
  co.time = p->time;
  co.theta1 = p->theta1;
  co.tau = p->tau;
  return co;
}

// Neumann boundary integrals of Assemble2D (src/AssembleRoutines/Assemble2D.C:7529-7570):
// line Gauss rule of degree 2*p, value g(comp, t) * hE * w times the test values.
void neumann_2d(const pmx_space* s, const pmx_cd_params* p, std::vector<double>& out) {
  const pmx::FESpace& fs = s->fs;
  const pmx::Mesh& m = *fs.mesh;
  out.assign(fs.n_dof, 0.0);
  if (m.dim != 2 || p->example != PMX_EX_SINE_LAPLACE_2D) return;
  const int npts = (fs.el.order == 1) ? 2 : 3;   // QFLineFromDegree[2p]
  double zx[5], zw[5];
  pmx::gauss_legendre(npts, zx, zw);
  std::vector<double> v(fs.ndof_cell);
  for (int c = 0; c < fs.n_cells; ++c) {
    for (int j = 0; j < m.n_faces; ++j) {
      const int comp = pmx::joint_boundary_component(m, fs.level, c, j);
      if (comp < 0) continue;
      int vv[2];
      pmx::joint_vertices(m.shape, j, vv);
      const double* a = m.vertex(fs.level, c, vv[0]);
      const double* b = m.vertex(fs.level, c, vv[1]);
      const double t0 = pmx::boundary_param_2d(comp, a[0], a[1]);
      const double t1 = pmx::boundary_param_2d(comp, b[0], b[1]);
      const int cond = s->bc2d ? s->bc2d(comp, 0.5 * (t0 + t1), s->user)
                               : builtin_cond2d(s->builtin_bc, comp, 0.5 * (t0 + t1));
      if (cond != pmx::BC_NEUMANN) continue;
      const pmx::AffineMap am = pmx::affine_map(m, fs.level, c);
      const double hE = std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1])) / 2;
      for (int l = 0; l < npts; ++l) {
        const double t = t0 + 0.5 * (t1 - t0) * (zx[l] + 1);
        double x[2] = {a[0] + 0.5 * (zx[l] + 1) * (b[0] - a[0]), a[1] + 0.5 * (zx[l] + 1) * (b[1] - a[1])};
        // SineLaplace BoundValue on component 1: eps * du/dx (1, t)
        double ex[6];
        pmx::example_exact(p->example, 1.0, t, 0.0, ex);
        double g = (comp == 1) ? p->eps * ex[1] : 0.0;
        g *= hE * zw[l];
        double xi[2];
        pmx::map_to_ref(am, x, xi);
        fs.el.eval(xi, v.data(), nullptr, nullptr);
        for (int k = 0; k < fs.ndof_cell; ++k) {
          const int dof = fs.global_numbers[(size_t)c * fs.ndof_cell + k];
          if (dof < fs.active_bound) out[dof] += g * v[k];
        }
      }
    }
  }
}

}  // namespace

// ========================================================================
void pmx_ctx_release(pmx_ctx* ctx) {
  if (ctx->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  dfree(ctx->partial);
  dfree(ctx->scal);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" {

const char* pmx_last_error(void) { return g_last_error.c_str(); }
int pmx_version(void) { return 100; }

int pmx_ctx_create(int device, pmx_ctx** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    require(device >= 0 && device < n, "device index out of range");
    HIP_CHECK(hipSetDevice(device));
    auto c = std::make_unique<pmx_ctx>();
    c->device = device;
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->partial = dmalloc<double>(4096);
    c->scal = dmalloc<double>(4096);
    // the fixed error-norm scratch of this device (pmx_compute_errors_noalloc)
    const int se = pmx::errors_scratch_init(device);
    if (se != 0) throw PmxError(PMX_E_HIP, "hipGetSymbolAddress of the error-norm scratch failed (code " +
                                               std::to_string(se) + ")");
    *out = c.release();
  });
}

int pmx_ctx_destroy(pmx_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    pmx_ctx_release(ctx);
  });
}

int pmx_ctx_synchronize(pmx_ctx* ctx) {
  return guarded([&] {
    require(ctx != nullptr, "ctx is NULL");
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pmx_domain_create(int geometry, int n_refine, pmx_domain** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    require(geometry >= 0 && geometry <= 2, "unknown geometry");
    auto d = std::make_unique<pmx_domain>();
    d->mesh = pmx::build_mesh(geometry, n_refine);
    *out = d.release();
  });
}

int pmx_domain_create_structured_tri(int n_side, pmx_domain** out) {
  return guarded([&] {
    require(out != nullptr, "out is NULL");
    require(n_side >= 1 && n_side <= 4096, "n_side must be in [1,4096]");
    auto d = std::make_unique<pmx_domain>();
    d->mesh = pmx::build_structured_tri_mesh(n_side);
    *out = d.release();
  });
}

int pmx_domain_destroy(pmx_domain* d) {
  return guarded([&] { delete d; });
}

int pmx_domain_level_info(const pmx_domain* d, int level, int* n_cells, int* nv, int* dim) {
  return guarded([&] {
    require(d != nullptr, "domain is NULL");
    require(level >= 0 && level < (int)d->mesh.levels.size(), "level out of range");
    if (n_cells) *n_cells = d->mesh.levels[level].n_cells;
    if (nv) *nv = d->mesh.nv;
    if (dim) *dim = d->mesh.dim;
  });
}

int pmx_domain_get_vertices(const pmx_domain* d, int level, double* host_xyz) {
  return guarded([&] {
    require(d != nullptr && host_xyz != nullptr, "NULL argument");
    require(level >= 0 && level < (int)d->mesh.levels.size(), "level out of range");
    const auto& x = d->mesh.levels[level].xyz;
    std::memcpy(host_xyz, x.data(), x.size() * sizeof(double));
  });
}

int pmx_space_create(pmx_ctx* ctx, const pmx_domain* d, int level, int order, pmx_bc2d_fn bc2d,
                     pmx_bc3d_fn bc3d, void* user, int builtin_bc, pmx_space** out) {
  return guarded([&] {
    require(d && out, "NULL argument");
    require(builtin_bc == PMX_BC_ALL_DIRICHLET || builtin_bc == PMX_BC_SINE_LAPLACE, "unknown builtin bc");
    auto s = std::make_unique<pmx_space>();
    s->ctx = ctx;
    s->dom = d;
    s->builtin_bc = builtin_bc;
    s->bc2d = bc2d;
    s->bc3d = bc3d;
    s->user = user;
    pmx::BoundaryConditions bc;
    pmx_space* sp = s.get();
    bc.cond2d = [sp](int comp, double t) {
      return sp->bc2d ? sp->bc2d(comp, t, sp->user) : builtin_cond2d(sp->builtin_bc, comp, t);
    };
    bc.cond3d = [sp](double x, double y, double z) {
      return sp->bc3d ? sp->bc3d(x, y, z, sp->user) : (int)pmx::BC_DIRICHLET;
    };
    s->fs = pmx::build_space(d->mesh, level, order, bc);
    const pmx::FESpace& fs = s->fs;
    const pmx::Mesh& m = d->mesh;
    s->pattern = std::make_shared<pmx::CSRPattern>(pmx::build_structure(fs));
    if (!ctx) { *out = s.release(); return; }
    // device geometry
    std::vector<double> geo((size_t)fs.n_cells * pmx::kGeoStride, 0.0);
    for (int c = 0; c < fs.n_cells; ++c) {
      const pmx::AffineMap a = pmx::affine_map(m, level, c);
      double* g = &geo[(size_t)c * pmx::kGeoStride];
      for (int r = 0; r < 3; ++r) g[r] = (r < m.dim) ? a.x0[r] : 0.0;
      for (int k = 0; k < m.dim * m.dim; ++k) { g[3 + k] = a.J[k]; g[12 + k] = a.Jinv[k]; }
      g[21] = a.detJ;
      g[22] = pmx::cell_diameter(m, level, c);
    }
    hipStream_t st = ctx->stream;
    s->d_geo = upload(geo.data(), geo.size(), st);
    s->d_dofs = upload(fs.global_numbers.data(), fs.global_numbers.size(), st);
    std::vector<int> order_cells;
    pmx::color_cells(fs, order_cells, s->color_ptr);
    s->d_cells = upload(order_cells.data(), order_cells.size(), st);
    const int ndir = fs.n_dof - fs.active_bound;
    s->d_dof_xyz = upload(fs.dof_xyz.data(), (size_t)fs.n_dof * 3, st);
    s->d_dir_xyz = s->d_dof_xyz + (size_t)fs.active_bound * 3;
    (void)ndir;
    *out = s.release();
  });
}

int pmx_space_destroy(pmx_space* s) {
  return guarded([&] { delete s; });
}

int pmx_space_get_info(const pmx_space* s, pmx_space_info* out) {
  return guarded([&] {
    require(s && out, "NULL argument");
    const pmx::FESpace& fs = s->fs;
    out->n_cells = fs.n_cells;
    out->ndof_cell = fs.ndof_cell;
    out->n_dof = fs.n_dof;
    out->active_bound = fs.active_bound;
    out->n_inner = fs.n_inner;
    out->n_dirichlet = fs.n_dirichlet;
    for (int k = 0; k < 9; ++k) out->n_bound[k] = fs.n_bound[k];
    out->dim = fs.el.dim;
  });
}

int pmx_space_get_global_numbers(const pmx_space* s, int* host_out) {
  return guarded([&] {
    require(s && host_out, "NULL argument");
    std::memcpy(host_out, s->fs.global_numbers.data(), s->fs.global_numbers.size() * sizeof(int));
  });
}

int pmx_space_get_dof_coords(const pmx_space* s, double* host_out) {
  return guarded([&] {
    require(s && host_out, "NULL argument");
    std::memcpy(host_out, s->fs.dof_xyz.data(), s->fs.dof_xyz.size() * sizeof(double));
  });
}

int pmx_space_get_structure(const pmx_space* s, int* nnz, int* rp, int* col) {
  return guarded([&] {
    require(s && nnz, "NULL argument");
    *nnz = (int)s->pattern->col.size();
    if (rp) std::memcpy(rp, s->pattern->row_ptr.data(), s->pattern->row_ptr.size() * sizeof(int));
    if (col) std::memcpy(col, s->pattern->col.data(), s->pattern->col.size() * sizeof(int));
  });
}

int pmx_space_get_prolongation(const pmx_space* coarse, const pmx_space* fine, int* nnz, int* rp,
                               int* col, double* val) {
  return guarded([&] {
    require(coarse && fine && nnz, "NULL argument");
    require(coarse->dom == fine->dom && fine->fs.level == coarse->fs.level + 1,
            "spaces must be consecutive levels of one domain");
    pmx::CSRPattern P;
    std::vector<double> v;
    pmx::build_prolongation(coarse->fs, fine->fs, P, v);
    *nnz = (int)P.col.size();
    if (rp) std::memcpy(rp, P.row_ptr.data(), P.row_ptr.size() * sizeof(int));
    if (col) std::memcpy(col, P.col.data(), P.col.size() * sizeof(int));
    if (val) std::memcpy(val, v.data(), v.size() * sizeof(double));
  });
}

int pmx_matrix_create(pmx_ctx* ctx, const pmx_space* s, pmx_matrix** out) {
  return guarded([&] {
    require(ctx && s && out, "NULL argument");
    require(s->ctx == ctx, "space was created without this device context");
    auto M = std::make_unique<pmx_matrix>();
    M->ctx = ctx;
    M->space = s;
    M->pattern = s->pattern;
    const pmx::CSRPattern& p = *M->pattern;
    hipStream_t st = ctx->stream;
    M->A.n = p.n_rows;
    M->A.m = p.n_cols;
    M->A.nnz = (int)p.col.size();
    M->A.active = p.active_bound;
    M->A.lpr = auto_lanes(M->A.nnz, M->A.n);
    M->A.row_ptr = upload(p.row_ptr.data(), p.row_ptr.size(), st);
    M->A.col = upload(p.col.data(), p.col.size(), st);
    M->A.val = dmalloc<double>(p.col.size());
    HIP_CHECK(hipMemsetAsync(M->A.val, 0, p.col.size() * sizeof(double), st));
    attach_sell(M.get(), st);
    HIP_CHECK(hipStreamSynchronize(st));
    *out = M.release();
  });
}

int pmx_matrix_create_csr(pmx_ctx* ctx, int n, int nnz, const int* rp, const int* col,
                          const double* val, int active_bound, pmx_matrix** out) {
  return guarded([&] {
    require(ctx && rp && col && val && out, "NULL argument");
    require(n >= 1 && nnz >= 0 && active_bound >= 0 && active_bound <= n, "invalid sizes");
    require(rp[0] == 0 && rp[n] == nnz, "row pointer inconsistent with nnz");
    for (int r = 0; r < n; ++r) {
      require(rp[r + 1] >= rp[r], "row pointer not monotone");
      for (int k = rp[r]; k < rp[r + 1]; ++k) {
        require(col[k] >= 0 && col[k] < n, "column index out of range");
        require(k == rp[r] || col[k] > col[k - 1], "columns must be sorted and unique per row");
      }
    }
    auto pat = std::make_shared<pmx::CSRPattern>();
    pat->n_rows = pat->n_cols = n;
    pat->active_bound = active_bound;
    pat->row_ptr.assign(rp, rp + n + 1);
    pat->col.assign(col, col + nnz);
    auto M = std::make_unique<pmx_matrix>();
    M->ctx = ctx;
    M->space = nullptr;
    M->pattern = pat;
    hipStream_t st = ctx->stream;
    M->A.n = M->A.m = n;
    M->A.nnz = nnz;
    M->A.active = active_bound;
    M->A.lpr = auto_lanes(nnz, n);
    M->A.row_ptr = upload(pat->row_ptr.data(), pat->row_ptr.size(), st);
    M->A.col = upload(pat->col.data(), pat->col.size(), st);
    M->A.val = upload(val, (size_t)nnz, st);
    attach_sell(M.get(), st);
    HIP_CHECK(hipStreamSynchronize(st));
    *out = M.release();
  });
}

int pmx_matrix_destroy(pmx_matrix* A) {
  return guarded([&] { delete A; });
}

int pmx_matrix_get_info(const pmx_matrix* A, int* n_rows, int* nnz, int* active_bound) {
  return guarded([&] {
    require(A != nullptr, "NULL matrix");
    if (n_rows) *n_rows = A->A.n;
    if (nnz) *nnz = A->A.nnz;
    if (active_bound) *active_bound = A->A.active;
  });
}

int pmx_matrix_get_structure(const pmx_matrix* A, int* rp, int* col) {
  return guarded([&] {
    require(A != nullptr, "NULL matrix");
    if (rp) std::memcpy(rp, A->pattern->row_ptr.data(), A->pattern->row_ptr.size() * sizeof(int));
    if (col) std::memcpy(col, A->pattern->col.data(), A->pattern->col.size() * sizeof(int));
  });
}

int pmx_matrix_get_entries(pmx_ctx* ctx, const pmx_matrix* A, double* host_values) {
  return guarded([&] {
    require(ctx && A && host_values, "NULL argument");
    HIP_CHECK(hipMemcpyAsync(host_values, A->A.val, (size_t)A->A.nnz * sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pmx_matrix_set_entries(pmx_ctx* ctx, pmx_matrix* A, const double* host_values) {
  return guarded([&] {
    require(ctx && A && host_values, "NULL argument");
    HIP_CHECK(hipMemcpyAsync(A->A.val, host_values, (size_t)A->A.nnz * sizeof(double),
                             hipMemcpyHostToDevice, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    A->sell_valid = false;
  });
}

int pmx_matrix_entries_device(pmx_matrix* A, double** dev_values) {
  return guarded([&] {
    require(A && dev_values, "NULL argument");
    *dev_values = A->A.val;
    A->sell_valid = false;   // the caller may write the values
  });
}

int pmx_matrix_reset(pmx_ctx* ctx, pmx_matrix* A) {
  return guarded([&] {
    require(ctx && A, "NULL argument");
    pmx::vec_fill(A->A.nnz, 0.0, A->A.val, ctx->stream);
    A->sell_valid = false;
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_matrix_set_spmv_lanes(pmx_matrix* A, int lanes) {
  return guarded([&] {
    require(A != nullptr, "NULL matrix");
    require(lanes == 0 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64,
            "lanes must be 0 (SELL-64 kernel), 4, 8, 16, 32 or 64");
    A->A.lpr = lanes;
  });
}

int pmx_matrix_apply_scaled_add(pmx_ctx* ctx, const pmx_matrix* A, const double* x, double* y,
                                double a) {
  return guarded([&] {
    require(ctx && A && x && y, "NULL argument");
    ensure_sell_values(const_cast<pmx_matrix*>(A), ctx->stream);
    pmx::fe_apply_scaled_add(A->A, x, y, a, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_matrix_apply(pmx_ctx* ctx, const pmx_matrix* A, const double* x, double* y) {
  return guarded([&] {
    require(ctx && A && x && y, "NULL argument");
    require(x != y, "x and y must not alias");
    ensure_sell_values(const_cast<pmx_matrix*>(A), ctx->stream);
    pmx::fe_apply(A->A, x, y, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_matrix_time_kernel(pmx_ctx* ctx, const pmx_matrix* A, int kernel, int reps, double* avg_ms) {
  return guarded([&] {
    require(ctx && A && avg_ms && reps >= 1, "invalid argument");
    require(kernel >= 0 && kernel <= 3, "unknown kernel");
    hipStream_t st = ctx->stream;
    const int n = A->A.n;
    ensure_sell_values(const_cast<pmx_matrix*>(A), st);
    pmx::DevCSR M = A->A;
    uint16_t* sb = nullptr;
    double *x = nullptr, *y = nullptr, *dinv = nullptr;
    float *x32 = nullptr, *y32 = nullptr, *d32 = nullptr, *rs = nullptr;
    uint16_t* vb = nullptr;
    x = dmalloc<double>(n); y = dmalloc<double>(n); dinv = dmalloc<double>(n);
    pmx::vec_fill(n, 1.0, x, st);
    pmx::vec_fill(n, 1.0, dinv, st);
    if (kernel >= 2) {
      x32 = dmalloc<float>(n); y32 = dmalloc<float>(n); d32 = dmalloc<float>(n); rs = dmalloc<float>(n);
      vb = dmalloc<uint16_t>(M.nnz);
      pmx::csr_to_bf16(M, vb, rs, st);
      M.val_bf16 = vb;
      M.rscale = rs;
      if (M.sell.total) {
        sb = dmalloc<uint16_t>(M.sell.total);
        pmx::sell_fill_bf16(M, sb, st);
        M.sell.val_bf16 = sb;
      }
      pmx::vec_fill_f32(n, 1.f, x32, st);
      pmx::vec_fill_f32(n, 1.f, d32, st);
    }
    auto launch = [&]() {
      switch (kernel) {
        case 0: pmx::fe_apply(M, x, y, st); break;
        case 1: pmx::fe_jacobi(M, x, x, dinv, y, 0.8, st); break;
        case 2: pmx::fe_apply_scaled_add_bf16(M, x32, y32, 1.0, st); break;
        case 3: pmx::fe_jacobi_bf16(M, x32, x32, d32, y32, 0.8, st); break;
      }
    };
    launch();
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch();
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipGetLastError());
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(x); dfree(y); dfree(dinv); dfree(x32); dfree(y32); dfree(d32); dfree(rs); dfree(vb); dfree(sb);
  });
}

int pmx_assembly_n_quad(const pmx_space* s, int hexa_quad_rule, int* n_quad) {
  return guarded([&] {
    require(s && n_quad, "NULL argument");
    *n_quad = pmx::assembly_rule(s->fs.el, s->fs.el.shape == pmx::SHAPE_HEXA ? hexa_quad_rule : 0).n;
  });
}

int pmx_assembly_quad_points(const pmx_space* s, int hexa_quad_rule, double* host_xyz) {
  return guarded([&] {
    require(s && host_xyz, "NULL argument");
    const pmx::FESpace& fs = s->fs;
    const pmx::QuadRule q = pmx::assembly_rule(fs.el, fs.el.shape == pmx::SHAPE_HEXA ? hexa_quad_rule : 0);
    for (int c = 0; c < fs.n_cells; ++c) {
      const pmx::AffineMap a = pmx::affine_map(*fs.mesh, fs.level, c);
      for (int k = 0; k < q.n; ++k) {
        double x[3] = {0, 0, 0};
        pmx::map_to_phys(a, &q.xi[k * q.dim], x);
        for (int d = 0; d < 3; ++d) host_xyz[((size_t)c * q.n + k) * 3 + d] = x[d];
      }
    }
  });
}

// Incidence slots of the Q2 row gather, built once per space: for every
// active row the (cell, local DOF) pairs in increasing cell order, the slot of
// each pair and the row positions of each such cell's 27 DOFs.  A space whose
// rows are too long or where too many cells meet at a DOF keeps the
// quadrature kernel.
bool q2_gather_ready(pmx_space* s, const pmx_matrix* A, hipStream_t st) {
  if (s->q2_gather_state) return s->q2_gather_state > 0;
  s->q2_gather_state = -1;
  const pmx::FESpace& fs = s->fs;
  const pmx::CSRPattern& pat = *s->pattern;
  const int na = fs.active_bound;
  for (int r = 0; r < na; ++r)
    if (pat.row_ptr[r + 1] - pat.row_ptr[r] > pmx::kQ2MaxRow) return false;
  std::vector<int> ptr(na + 1, 0);
  const std::vector<int>& gn = fs.global_numbers;
  for (size_t k = 0; k < gn.size(); ++k)
    if (gn[k] < na) ++ptr[gn[k] + 1];
  for (int r = 0; r < na; ++r) {
    if (ptr[r + 1] > pmx::kQ2MaxInc) return false;
    ptr[r + 1] += ptr[r];
  }
  std::vector<int> inc(ptr[na]), slot(gn.size(), -1), fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < fs.n_cells; ++e)
    for (int i = 0; i < 27; ++i) {
      const int r = gn[(size_t)e * 27 + i];
      if (r >= na) continue;
      slot[(size_t)e * 27 + i] = fill[r];
      inc[fill[r]++] = 27 * e + i;
    }
  s->d_q2_inc_ptr = upload(ptr.data(), ptr.size(), st);
  s->d_q2_slot = upload(slot.data(), slot.size(), st);
  s->d_q2_off = dmalloc<uint8_t>((size_t)inc.size() * 32);
  pmx::Q2Gather g;
  g.n_active = na;
  g.n_inc = (int)inc.size();
  g.inc_ptr = s->d_q2_inc_ptr;
  g.slot = s->d_q2_slot;
  g.off = s->d_q2_off;
  int* d_inc = upload(inc.data(), inc.size(), st);
  int* d_bad = dmalloc<int>(1);
  HIP_CHECK(hipMemsetAsync(d_bad, 0, sizeof(int), st));
  pmx::q2_gather_offsets(g, d_inc, s->d_dofs, A->A, s->d_q2_off, d_bad, st);
  HIP_CHECK(hipGetLastError());
  int bad = 0;
  HIP_CHECK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  dfree(d_bad);
  dfree(d_inc);
  if (bad) {
    dfree(s->d_q2_inc_ptr); dfree(s->d_q2_slot); dfree(s->d_q2_off);
    return false;
  }
  s->q2g = g;
  s->q2_gather_state = 1;
  return true;
}

// element GEMM + row gather (Q2 hexahedra, constant coefficients per cell)
void assemble_cd_q2_gemm(pmx_space* s, RefTable& t, const pmx::AssemblyData& ad, const pmx::CDCoeffs& co,
                         bool fp32, bool overwrite, pmx_matrix* A, double* dev_rhs, hipStream_t st) {
  if (!t.d_R64) {
    std::vector<double> R((size_t)pmx::kQ2Terms * pmx::kQ2N);
    pmx::q2_reference_matrices(t.h_ref.data(), t.nq, R.data());
    std::vector<float> R32(R.begin(), R.end());
    t.d_R64 = upload(R.data(), R.size(), st);
    t.d_R32 = upload(R32.data(), R32.size(), st);
  }
  const pmx::Q2Gather& g = s->q2g;
  const size_t ne = (size_t)g.n_inc * 32;
  if (!s->d_q2_fe) s->d_q2_fe = dmalloc<double>(std::max(1, g.n_inc));
  if (fp32) {
    if (!s->d_q2_E32) s->d_q2_E32 = dmalloc<float>(std::max<size_t>(1, ne));
    pmx::q2_element_gemm(ad, co, g, t.d_R32, s->d_q2_E32, s->d_q2_fe, st);
    pmx::q2_row_gather(g, s->d_q2_E32, s->d_q2_fe, A->A, dev_rhs, overwrite, st);
  } else {
    if (!s->d_q2_E64) s->d_q2_E64 = dmalloc<double>(std::max<size_t>(1, ne));
    pmx::q2_element_gemm(ad, co, g, t.d_R64, s->d_q2_E64, s->d_q2_fe, st);
    pmx::q2_row_gather(g, s->d_q2_E64, s->d_q2_fe, A->A, dev_rhs, overwrite, st);
  }
}

int pmx_assemble_cd(pmx_ctx* ctx, const pmx_space* s_, const pmx_cd_params* p, pmx_matrix* A,
                    double* dev_rhs) {
  return guarded([&] {
    require(ctx && s_ && p && A && dev_rhs, "NULL argument");
    require(A->space == s_, "matrix was not created on this space");
    require(s_->ctx == ctx, "space has no device data for this context");
    require(p->disc == PMX_DISC_GALERKIN || p->disc == PMX_DISC_SUPG || p->disc == PMX_DISC_GLS,
            "unknown discretisation");
    require(p->example >= 0 && p->example <= PMX_EX_TCD_LINEAR, "unknown example");
    require(p->assembly_variant == 0 || p->assembly_variant == 1, "assembly_variant must be 0 or 1");
    pmx_space* s = const_cast<pmx_space*>(s_);
    hipStream_t st = ctx->stream;
    A->sell_valid = false;
    RefTable& t = const_cast<RefTable&>(ref_table(s, p->hexa_quad_rule));
    pmx::AssemblyData ad = assembly_data(s, t);
    const pmx::CDCoeffs co = cd_coeffs(p);
    const bool gemm = ad.dim == 3 && ad.ndof == 27 && p->assembly_variant == 0 && !co.table && co.supg != 2 &&
                      ad.nq <= 64 && q2_gather_ready(s, A, st);
    require(!p->fp32_assembly || gemm,
            "fp32_assembly needs the element-GEMM route: Q2 hexahedra, Galerkin or SUPG, no coefficient "
            "table, at most 64 quadrature points, assembly_variant 0");
    if (p->overwrite && !gemm) {
      pmx::vec_fill(A->A.nnz, 0.0, A->A.val, st);
      pmx::vec_fill(s->fs.n_dof, 0.0, dev_rhs, st);
    }
    if (gemm) {
      assemble_cd_q2_gemm(s, t, ad, co, p->fp32_assembly != 0, p->overwrite != 0, A, dev_rhs, st);
    } else if (ad.dim == 3 && ad.ndof == 27) {
      // the matrix-core kernel scatters through a position table of the
      // space's pattern (built once, 4 KB per cell)
      if (!s->d_q2_pos) {
        s->d_q2_pos = dmalloc<int>((size_t)ad.n_cells * pmx::kQ2PosPerCell);
        pmx::q2_scatter_positions(ad, A->A, s->d_q2_pos, st);
        HIP_CHECK(hipGetLastError());
      }
      ad.pos = s->d_q2_pos;
    }
    if (!gemm)
      for (size_t k = 0; k + 1 < s->color_ptr.size(); ++k) {
        const int n = s->color_ptr[k + 1] - s->color_ptr[k];
        pmx::assemble_cd(ad, co, s->d_cells + s->color_ptr[k], n, A->A, dev_rhs, st);
      }
    HIP_CHECK(hipGetLastError());
    const pmx::FESpace& fs = s->fs;
    const int ndir = fs.n_dof - fs.active_bound;
    if (ndir > 0) {
      if (p->host_dirichlet) {
        HIP_CHECK(hipMemcpyAsync(dev_rhs + fs.active_bound, p->host_dirichlet, ndir * sizeof(double),
                                 hipMemcpyHostToDevice, st));
      } else {
        pmx::dirichlet_rhs_example(p->example, ndir, s->d_dir_xyz, dev_rhs + fs.active_bound, co, st);
      }
    }
    if (fs.el.dim == 2 && p->example == PMX_EX_SINE_LAPLACE_2D) {
      std::vector<double> neu;
      neumann_2d(s, p, neu);
      double* d = upload(neu.data(), neu.size(), st);
      pmx::vec_axpy(fs.n_dof, 1.0, nullptr, 1.0, d, dev_rhs, st);
      HIP_CHECK(hipStreamSynchronize(st));
      dfree(d);
    }
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_assemble_tcd(pmx_ctx* ctx, const pmx_space* s_, const pmx_cd_params* p, pmx_matrix* A,
                     pmx_matrix* M, double* dev_rhs) {
  return guarded([&] {
    require(ctx && s_ && p, "NULL argument");
    require(A || M || dev_rhs, "nothing to assemble");
    require(!A || A->space == s_, "stiffness matrix was not created on this space");
    require(!M || M->space == s_, "mass matrix was not created on this space");
    require(s_->ctx == ctx, "space has no device data for this context");
    require(p->disc == PMX_DISC_GALERKIN || p->disc == PMX_DISC_SUPG, "unknown discretisation");
    require(p->disc == PMX_DISC_GALERKIN || s_->fs.el.dim == 2,
            "SUPG for time-dependent problems is implemented in 2D only (as TCDStiffSUPG<2>)");
    require(p->example >= 0 && p->example <= PMX_EX_TCD_LINEAR, "unknown example");
    require(p->disc == PMX_DISC_GALERKIN || p->tau > 0, "SUPG needs the time step length tau");
    pmx_space* s = const_cast<pmx_space*>(s_);
    hipStream_t st = ctx->stream;
    const RefTable& t = ref_table(s, p->hexa_quad_rule);
    const pmx::AssemblyData ad = assembly_data(s, t);
    const pmx::CDCoeffs co = cd_coeffs(p);
    pmx::DevCSR none{};
    const pmx::DevCSR& dA = A ? A->A : none;
    const pmx::DevCSR& dM = M ? M->A : none;
    if (A) A->sell_valid = false;
    if (M) M->sell_valid = false;
    for (size_t k = 0; k + 1 < s->color_ptr.size(); ++k) {
      const int n = s->color_ptr[k + 1] - s->color_ptr[k];
      pmx::assemble_tcd(ad, co, s->d_cells + s->color_ptr[k], n, s->fs.active_bound, dA, dM, dev_rhs, st);
    }
    HIP_CHECK(hipGetLastError());
    const pmx::FESpace& fs = s->fs;
    const int ndir = fs.n_dof - fs.active_bound;
    if (dev_rhs && ndir > 0) {
      if (p->host_dirichlet) {
        HIP_CHECK(hipMemcpyAsync(dev_rhs + fs.active_bound, p->host_dirichlet, ndir * sizeof(double),
                                 hipMemcpyHostToDevice, st));
      } else {
        pmx::dirichlet_rhs_example(p->example, ndir, s->d_dir_xyz, dev_rhs + fs.active_bound, co, st);
      }
    }
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_interpolate(pmx_ctx* ctx, const pmx_space* s, const pmx_cd_params* p, double* dev_u) {
  return guarded([&] {
    require(ctx && s && p && dev_u, "NULL argument");
    require(s->ctx == ctx, "space has no device data for this context");
    require(p->example >= 0 && p->example <= PMX_EX_TCD_LINEAR, "unknown example");
    pmx::interpolate_example(p->example, p->time, s->fs.n_dof, s->d_dof_xyz, dev_u, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

namespace {
int active_entries(const pmx_matrix* A) { return A->pattern->row_ptr[A->pattern->active_bound]; }
void require_same_pattern(const pmx_matrix* A, const pmx_matrix* B) {
  require(A->pattern == B->pattern ||
              (A->pattern->row_ptr == B->pattern->row_ptr && A->pattern->col == B->pattern->col &&
               A->pattern->active_bound == B->pattern->active_bound),
          "matrices do not share one sparsity pattern");
}
}  // namespace

int pmx_matrix_scale_active(pmx_ctx* ctx, pmx_matrix* A, double factor) {
  return guarded([&] {
    require(ctx && A, "NULL argument");
    if (factor == 1.0) return;
    const int na = active_entries(A);
    if (factor == 0.0) pmx::vec_fill(na, 0.0, A->A.val, ctx->stream);
    else pmx::vec_scale(na, factor, nullptr, 0, A->A.val, A->A.val, ctx->stream);
    A->sell_valid = false;
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_matrix_add_scaled_active(pmx_ctx* ctx, pmx_matrix* A, const pmx_matrix* B, double factor) {
  return guarded([&] {
    require(ctx && A && B, "NULL argument");
    require_same_pattern(A, B);
    pmx::vec_axpy(active_entries(A), factor, nullptr, 1.0, B->A.val, A->A.val, ctx->stream);
    A->sell_valid = false;
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_matrix_multiply_active(pmx_ctx* ctx, const pmx_matrix* A, const double* x, double* y, double a) {
  return guarded([&] {
    require(ctx && A && x && y, "NULL argument");
    require(x != y, "x and y must not alias");
    ensure_sell_values(const_cast<pmx_matrix*>(A), ctx->stream);
    pmx::fe_multiply_active(A->A, x, y, a, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_vector_scale(pmx_ctx* ctx, int n, double a, double* x) {
  return guarded([&] {
    require(ctx && x && n >= 0, "invalid argument");
    pmx::vec_scale(n, a, nullptr, 0, x, x, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_vector_copy(pmx_ctx* ctx, int n, const double* x, double* y) {
  return guarded([&] {
    require(ctx && x && y && n >= 0, "invalid argument");
    if (n > 0 && x != y)
      HIP_CHECK(hipMemcpyAsync(y, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  });
}

// ------------------------------------------------------- time discretisation
namespace {
// factor of the stiffness matrix in the system matrix
// (TimeDiscretization::prepare_system_matrix, TimeDiscretizations.C:172-210)
double system_factor(int scheme, int step, double tau) {
  if (scheme == PMX_TIME_CRANK_NICOLSON) return 0.5 * tau;
  if (scheme == PMX_TIME_BDF2 && step >= 2) return tau * (2.0 / 3.0);
  return tau;
}
void require_scheme(int scheme, double tau) {
  require(scheme == PMX_TIME_BACKWARD_EULER || scheme == PMX_TIME_CRANK_NICOLSON || scheme == PMX_TIME_BDF2,
          "time stepping scheme is not supported");
  require(tau > 0, "time step length must be positive");
}
}  // namespace

int pmx_time_prepare_rhs(pmx_ctx* ctx, int scheme, int step, double tau, const pmx_matrix* A,
                         const pmx_matrix* M, double* rhs, const double* old_rhs, const double* u_m1,
                         const double* u_m2) {
  return guarded([&] {
    require(ctx && A && M && rhs && u_m1, "NULL argument");
    require_scheme(scheme, tau);
    require_same_pattern(A, M);
    hipStream_t st = ctx->stream;
    const int ab = A->pattern->active_bound;
    ensure_sell_values(const_cast<pmx_matrix*>(M), st);
    if (scheme == PMX_TIME_BDF2 && step >= 2) {
      require(u_m2, "BDF2 needs the solution two steps back");
      pmx::vec_scale(ab, (2.0 / 3.0) * tau, nullptr, 0, rhs, rhs, st);
      pmx::fe_multiply_active(M->A, u_m1, rhs, 4.0 / 3.0, st);
      pmx::fe_multiply_active(M->A, u_m2, rhs, -1.0 / 3.0, st);
    } else if (scheme == PMX_TIME_CRANK_NICOLSON) {
      require(old_rhs, "Crank-Nicolson needs the right-hand side of the previous step");
      ensure_sell_values(const_cast<pmx_matrix*>(A), st);
      pmx::vec_scale(ab, 0.5 * tau, nullptr, 0, rhs, rhs, st);
      pmx::vec_axpy(ab, 0.5 * tau, nullptr, 1.0, old_rhs, rhs, st);
      pmx::fe_multiply_active(M->A, u_m1, rhs, 1.0, st);
      pmx::fe_multiply_active(A->A, u_m1, rhs, -0.5 * tau, st);
    } else {
      pmx::vec_scale(ab, tau, nullptr, 0, rhs, rhs, st);
      pmx::fe_multiply_active(M->A, u_m1, rhs, 1.0, st);
    }
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_time_prepare_system(pmx_ctx* ctx, int scheme, int step, double tau, pmx_matrix* A,
                            const pmx_matrix* M) {
  return guarded([&] {
    require(ctx && A && M, "NULL argument");
    require_scheme(scheme, tau);
    require_same_pattern(A, M);
    const int na = active_entries(A);
    pmx::vec_scale(na, system_factor(scheme, step, tau), nullptr, 0, A->A.val, A->A.val, ctx->stream);
    pmx::vec_axpy(na, 1.0, nullptr, 1.0, M->A.val, A->A.val, ctx->stream);
    A->sell_valid = false;
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_time_reset_system(pmx_ctx* ctx, int scheme, int step, double tau, pmx_matrix* A,
                          const pmx_matrix* M) {
  return guarded([&] {
    require(ctx && A && M, "NULL argument");
    require_scheme(scheme, tau);
    require_same_pattern(A, M);
    const int na = active_entries(A);
    pmx::vec_axpy(na, -1.0, nullptr, 1.0, M->A.val, A->A.val, ctx->stream);
    pmx::vec_scale(na, 1.0 / system_factor(scheme, step, tau), nullptr, 0, A->A.val, A->A.val, ctx->stream);
    A->sell_valid = false;
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_compute_errors(pmx_ctx* ctx, const pmx_space* s_, const pmx_cd_params* p,
                       const double* dev_u, double* host_errors4) {
  return guarded([&] {
    require(ctx && s_ && p && dev_u && host_errors4, "NULL argument");
    require(s_->ctx == ctx, "space has no device data for this context");
    pmx_space* s = const_cast<pmx_space*>(s_);
    hipStream_t st = ctx->stream;
    const RefTable& t = ref_table(s, p->hexa_quad_rule);
    const pmx::AssemblyData ad = assembly_data(s, t);
    const pmx::CDCoeffs co = cd_coeffs(p);
    const int nb = (ad.n_cells + 255) / 256;
    double* partial = dmalloc<double>((size_t)nb * 4);
    pmx::cd_errors(ad, co, dev_u, partial, ctx->scal, st);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(host_errors4, ctx->scal, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    dfree(partial);
  });
}

int pmx_compute_errors_noalloc(pmx_ctx* ctx, const pmx_space* s_, const pmx_cd_params* p,
                               const double* dev_u, double* host_errors4) {
  return guarded([&] {
    require(ctx && s_ && p && dev_u && host_errors4, "NULL argument");
    require(s_->ctx == ctx, "space has no device data for this context");
    const int key = (s_->fs.el.shape == pmx::SHAPE_HEXA) ? p->hexa_quad_rule : 0;
    const auto it = s_->ref.find(key);
    if (it == s_->ref.end())
      throw PmxError(PMX_E_STATE, "the space's quadrature tables for this rule are not on the device yet: "
                                  "call pmx_assemble_cd, pmx_assemble_tcd or pmx_compute_errors on it once first");
    if (s_->fs.n_cells > 256 * pmx::kErrorsScratchBlocks)
      throw PmxError(PMX_E_ARG, "pmx_compute_errors_noalloc supports at most " +
                                    std::to_string(256 * pmx::kErrorsScratchBlocks) + " cells, the space has " +
                                    std::to_string(s_->fs.n_cells));
    double* scratch = pmx::errors_scratch(ctx->device);
    if (!scratch) throw PmxError(PMX_E_STATE, "error-norm scratch not resolved on this device");
    const pmx::AssemblyData ad = assembly_data(const_cast<pmx_space*>(s_), it->second);
    const pmx::CDCoeffs co = cd_coeffs(p);
    const double* out4 = nullptr;
    require(pmx::cd_errors_scratch(ad, co, dev_u, scratch, &out4, ctx->stream), "space too large for the scratch");
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(host_errors4, out4, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pmx_vector_dot(pmx_ctx* ctx, int n, const double* x, const double* y, double* out) {
  return guarded([&] {
    require(ctx && x && y && out && n >= 0, "invalid argument");
    pmx::vec_dot(n, x, y, ctx->partial, ctx->scal, ctx->stream);
    HIP_CHECK(hipMemcpyAsync(out, ctx->scal, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pmx_vector_norm(pmx_ctx* ctx, int n, const double* x, double* out) {
  return guarded([&] {
    require(ctx && x && out && n >= 0, "invalid argument");
    pmx::vec_norm(n, x, ctx->partial, ctx->scal, ctx->stream);
    HIP_CHECK(hipMemcpyAsync(out, ctx->scal, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int pmx_vector_add_scaled(pmx_ctx* ctx, int n, double a, const double* x, double* y) {
  return guarded([&] {
    require(ctx && x && y && n >= 0, "invalid argument");
    pmx::vec_axpy(n, a, nullptr, 1.0, x, y, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
