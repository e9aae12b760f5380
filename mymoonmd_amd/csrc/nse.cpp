// Oseen / Navier-Stokes saddle-point systems on hexahedra (Q2 velocity, Q1
// pressure) with the NSTYPE-1 block layout of BlockFEMatrix::NSE3D_Type1
// (src/Matrix/BlockFEMatrix.C:471-720):
//     ( A   0   0   B1^T )
//     ( 0   A   0   B2^T )      A: velocity block shared by the 3 components,
//     ( 0   0   A   B3^T )      Bk: pressure rows x velocity columns,
//     ( B1  B2  B3  0    )      Bk^T applied as the transposes of Bk.
// Vectors are BlockVectors [u1 | u2 | u3 | p], every block in its space's
// numbering (velocity Dirichlet DOFs last; pressure DOFs all active).
//
// The C ABI pmx_nse_* of include/pmx.h lives here: the system of one level
// (spaces, structures, assembly, block operator, combined matrix, errors,
// pressure projection).  The multigrid solver over several levels is in
// nse_solver.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "hip/examples.hpp"
#include "hip/launch.hpp"
#include "host/fe.hpp"
#include "host/mesh.hpp"
#include "host/space.hpp"
#include "internal.hpp"
#include "nse_internal.hpp"
#include "pmx.h"

using namespace pmx_internal;

pmx_nse::~pmx_nse() {
  delete A;
  for (void* p : std::initializer_list<void*>{d_geo, d_vdofs, d_pdofs, d_cells, d_vxyz, d_ref, d_Brp, d_Bcol, d_B1,
                                              d_B2, d_B3, d_BTrp, d_BTcol, d_BTsrc, d_BT1, d_BT2, d_BT3, d_Crp,
                                              d_Ccol, d_Cval, d_Ckind, d_Cpos, d_partial, d_pref, d_pgeo})
    dfree(p);
  for (auto& kv : err_ref) dfree(kv.second);
  for (auto& kv : restrict_from) {
    dfree(kv.second.row_ptr);
    dfree(kv.second.col);
    dfree(kv.second.val);
  }
}

namespace pmx_internal {

pmx::DevCSR nse_bt(const pmx_nse* s, int k) {
  pmx::DevCSR T{};
  T.n = s->n_u;
  T.m = s->n_p;
  T.nnz = (int)s->BTp.col.size();
  T.active = s->n_u_active;
  T.lpr = 8;
  T.row_ptr = s->d_BTrp;
  T.col = s->d_BTcol;
  T.val = (k == 0) ? s->d_BT1 : (k == 1) ? s->d_BT2 : s->d_BT3;
  return T;
}

pmx::DevCSR nse_b(const pmx_nse* s, int k) {
  pmx::DevCSR B{};
  B.n = s->n_p;
  B.m = s->n_u;
  B.nnz = (int)s->Bp.col.size();
  B.active = s->n_p;
  B.lpr = 32;
  B.row_ptr = s->d_Brp;
  B.col = s->d_Bcol;
  B.val = (k == 0) ? s->d_B1 : (k == 1) ? s->d_B2 : s->d_B3;
  return B;
}

pmx::DevCSR nse_combined(const pmx_nse* s) {
  pmx::DevCSR C{};
  C.n = C.m = s->n_comb;
  C.nnz = (int)s->Cp.col.size();
  C.active = s->n_comb;
  // 16 lanes per row: the rows are long (83 entries on average at 32^3) and
  // few; one row per lane (SELL) leaves too little parallelism per wave, 8 or
  // 32 lanes per row measured slower on MI355X
  C.lpr = 16;
  C.row_ptr = s->d_Crp;
  C.col = s->d_Ccol;
  C.val = s->d_Cval;
  return C;
}

// y = K x (+ y * beta_add): BlockFEMatrix::apply_scaled_add with scale a
void nse_apply_scaled_add(const pmx_nse* s, const double* x, double* y, double a, hipStream_t st) {
  const int nu = s->n_u;
  ensure_sell_values(s->A, st);
  // A on the three velocity components in one pass over A
  pmx::fe_apply_scaled_add3(s->A->A, x, y, (size_t)nu, a, st);
  // B_k^T p on the active velocity rows (FEMatrix::multiplyTransposedActive),
  // then B_1 u_1 + B_2 u_2 + B_3 u_3 on the pressure rows, each with one
  // pass over its pattern
  pmx::csr_bt3(nse_bt(s, 0), s->d_BT2, s->d_BT3, x + (size_t)3 * nu, y, (size_t)nu, a, s->n_u_active, st);
  pmx::csr_b3(nse_b(s, 0), s->d_B2, s->d_B3, x, (size_t)nu, y + (size_t)3 * nu, a, s->n_p, st);
}

void nse_refresh(pmx_nse* s, hipStream_t st) {
  // B^T values from B, combined matrix values from the blocks
  const long long nbt = (long long)s->BTp.col.size();
  pmx::sell_gather(nbt, s->d_BTsrc, s->d_B1, s->d_BT1, nullptr, st);
  pmx::sell_gather(nbt, s->d_BTsrc, s->d_B2, s->d_BT2, nullptr, st);
  pmx::sell_gather(nbt, s->d_BTsrc, s->d_B3, s->d_BT3, nullptr, st);
  pmx::combined_values((long long)s->Cp.col.size(), s->d_Ckind, s->d_Cpos, s->A->A.val, s->d_B1, s->d_B2, s->d_B3,
                       s->d_Cval, st);
  s->values_ready = true;
}

}  // namespace pmx_internal

namespace {

// reference table of the velocity/pressure pair at the quadrature points of
// the velocity rule: w, xi[3], phi[27], dphi[3][27], psi[8] (a 4-DOF
// pressure element fills psi[0..3], the rest stays 0)
std::vector<double> nse_ref_table(const pmx::QuadRule& rule, const pmx::Element& ep) {
  const pmx::Element ev = pmx::make_element(pmx::SHAPE_HEXA, 2);
  const int rs = 1 + 3 + 27 + 81 + 8;
  std::vector<double> t((size_t)rs * rule.n, 0.0);
  std::vector<double> d2(6 * 27), pv(8), pg(24), pd2(48);
  for (int q = 0; q < rule.n; ++q) {
    double* r = &t[(size_t)q * rs];
    r[0] = rule.w[q];
    for (int a = 0; a < 3; ++a) r[1 + a] = rule.xi[3 * q + a];
    ev.eval(&rule.xi[3 * q], r + 4, r + 4 + 27, d2.data());
    ep.eval(&rule.xi[3 * q], pv.data(), pg.data(), pd2.data());
    for (int j = 0; j < ep.ndof; ++j) r[4 + 27 + 81 + j] = pv[j];
  }
  return t;
}

// example id of component k (0..2 velocity, 3 pressure)
int nse_component(int example, int k) {
  return (example == PMX_EX_NSE_U0P0 ? pmx::EX_NSE_U0P0_U1 : pmx::EX_NSE_U2P1_U1) + k;
}

// test_u_0_p_0.h on_front_face_interior: the face x = 1 without its boundary
bool front_face_interior(double x, double y, double z) {
  const double tol = 1e-10;
  return std::fabs(1 - x) < tol && !(std::fabs(y - 1) < tol) && !(std::fabs(y) < tol) &&
         !(std::fabs(z - 1) < tol) && !(std::fabs(z) < tol);
}

// scalar reference table (w, xi, phi, dphi, d2phi) for the error kernel
double* scalar_ref(pmx_nse* s, const pmx::FESpace& fs, int* nq) {
  const int key = fs.el.disc ? -fs.el.order : fs.el.order;
  auto it = s->err_ref.find(key);
  const pmx::QuadRule rule = pmx::assembly_rule(fs.el, 0);
  *nq = rule.n;
  if (it != s->err_ref.end()) return it->second;
  const int nd = fs.el.ndof, rs = 1 + 3 + nd * (1 + 3 + 6);
  std::vector<double> t((size_t)rs * rule.n, 0.0);
  for (int q = 0; q < rule.n; ++q) {
    double* r = &t[(size_t)q * rs];
    r[0] = rule.w[q];
    for (int a = 0; a < 3; ++a) r[1 + a] = rule.xi[3 * q + a];
    fs.el.eval(&rule.xi[3 * q], r + 4, r + 4 + nd, r + 4 + 4 * nd);
  }
  double* d = upload(t.data(), t.size(), s->ctx->stream);
  s->err_ref[key] = d;
  return d;
}

}  // namespace

extern "C" {

int pmx_nse_create(pmx_ctx* ctx, const pmx_domain* d, int level, pmx_nse** out) {
  return pmx_nse_create_pair(ctx, d, level, PMX_NSE_PRESSURE_Q1, PMX_EX_NSE_U2P1, out);
}

int pmx_nse_create_pair(pmx_ctx* ctx, const pmx_domain* d, int level, int pressure_space, int example,
                        pmx_nse** out) {
  return guarded([&] {
    require(ctx && d && out, "NULL argument");
    require(d->mesh.shape == pmx::SHAPE_HEXA, "the saddle-point path is implemented for hexahedra");
    require(level >= 0 && level < (int)d->mesh.levels.size(), "level out of range");
    require(pressure_space == PMX_NSE_PRESSURE_Q1 || pressure_space == PMX_NSE_PRESSURE_P1DISC,
            "pressure space must be PMX_NSE_PRESSURE_Q1 or PMX_NSE_PRESSURE_P1DISC");
    require(example == PMX_EX_NSE_U2P1 || example == PMX_EX_NSE_U0P0, "unknown Navier-Stokes example");
    auto s = std::make_unique<pmx_nse>();
    s->ctx = ctx;
    s->dom = d;
    s->level = level;
    s->pressure_space = pressure_space;
    s->example = example;
    pmx::BoundaryConditions vbc, pbc;
    vbc.cond2d = [](int, double) { return (int)pmx::BC_DIRICHLET; };
    if (example == PMX_EX_NSE_U0P0)
      vbc.cond3d = [](double x, double y, double z) {
        return (int)(front_face_interior(x, y, z) ? pmx::BC_NEUMANN : pmx::BC_DIRICHLET);
      };
    else
      vbc.cond3d = [](double, double, double) { return (int)pmx::BC_DIRICHLET; };
    pbc.cond2d = [](int, double) { return (int)pmx::BC_NEUMANN; };
    pbc.cond3d = [](double, double, double) { return (int)pmx::BC_NEUMANN; };
    s->vs = pmx::build_space(d->mesh, level, 2, vbc);
    s->ps = (pressure_space == PMX_NSE_PRESSURE_P1DISC) ? pmx::build_space_p1disc(d->mesh, level)
                                                        : pmx::build_space(d->mesh, level, 1, pbc);
    s->np_cell = s->ps.ndof_cell;
    // BlockFEMatrix.C:1869-1920: an enclosed flow has no Neumann or Robin velocity DOF
    s->enclosed = s->vs.n_bound[pmx::BC_NEUMANN - 1] == 0 && s->vs.n_bound[pmx::BC_ROBIN - 1] == 0;
    s->n_u = s->vs.n_dof;
    s->n_u_active = s->vs.active_bound;
    s->n_p = s->ps.n_dof;
    s->n_comb = 3 * s->n_u + s->n_p;
    auto Ap = std::make_shared<pmx::CSRPattern>(pmx::build_structure(s->vs));
    s->Bp = pmx::build_rect_structure(s->ps, s->vs);
    // B^T pattern and the B position of every B^T entry
    {
      std::vector<double> pos(s->Bp.col.size());
      for (size_t k = 0; k < pos.size(); ++k) pos[k] = (double)k;
      std::vector<double> tpos;
      pmx::transpose_csr(s->Bp, pos, s->BTp, tpos);
      s->BTsrc.resize(tpos.size());
      for (size_t k = 0; k < tpos.size(); ++k) s->BTsrc[k] = (int)tpos[k];
    }
    // combined matrix (BlockFEMatrix::get_combined_matrix): Dirichlet velocity
    // rows are unit rows; for an enclosed flow the first pressure row becomes
    // a unit row (pressure projection, BlockFEMatrix.C:1237-1300)
    {
      const int nu = s->n_u, na = s->n_u_active, np = s->n_p;
      const pmx::CSRPattern& A = *Ap;
      std::vector<int>& rp = s->Cp.row_ptr;
      std::vector<int>& col = s->Cp.col;
      rp.assign(s->n_comb + 1, 0);
      for (int k = 0; k < 3; ++k) {
        for (int r = 0; r < nu; ++r) {
          const int row = k * nu + r;
          if (r >= na) {
            col.push_back(row);
            s->Ckind.push_back(1);
            s->Cpos.push_back(0);
          } else {
            for (int q = A.row_ptr[r]; q < A.row_ptr[r + 1]; ++q) {
              col.push_back(k * nu + A.col[q]);
              s->Ckind.push_back(2);
              s->Cpos.push_back(q);
            }
            for (int q = s->BTp.row_ptr[r]; q < s->BTp.row_ptr[r + 1]; ++q) {
              col.push_back(3 * nu + s->BTp.col[q]);
              s->Ckind.push_back((uint8_t)(3 + k));
              s->Cpos.push_back(s->BTsrc[q]);
            }
          }
          rp[row + 1] = (int)col.size();
        }
      }
      for (int r = 0; r < np; ++r) {
        const int row = 3 * nu + r;
        if (r == 0 && s->enclosed) {
          col.push_back(row);
          s->Ckind.push_back(1);
          s->Cpos.push_back(0);
        } else {
          for (int k = 0; k < 3; ++k)
            for (int q = s->Bp.row_ptr[r]; q < s->Bp.row_ptr[r + 1]; ++q) {
              col.push_back(k * nu + s->Bp.col[q]);
              s->Ckind.push_back((uint8_t)(3 + k));
              s->Cpos.push_back(q);
            }
        }
        rp[row + 1] = (int)col.size();
      }
      s->Cp.n_rows = s->Cp.n_cols = s->n_comb;
      s->Cp.active_bound = s->n_comb;
    }
    hipStream_t st = ctx->stream;
    // velocity block with its SELL copy
    s->A = new pmx_matrix();
    s->A->ctx = ctx;
    s->A->pattern = Ap;
    s->A->A.n = s->A->A.m = s->n_u;
    s->A->A.nnz = (int)Ap->col.size();
    s->A->A.active = s->n_u_active;
    s->A->A.row_ptr = upload(Ap->row_ptr.data(), Ap->row_ptr.size(), st);
    s->A->A.col = upload(Ap->col.data(), Ap->col.size(), st);
    s->A->A.val = dmalloc<double>(Ap->col.size());
    attach_sell(s->A, st);
    // geometry and DOF maps (cells in colour order for the scatter)
    const pmx::Mesh& m = d->mesh;
    const int nc = s->vs.n_cells;
    std::vector<double> geo((size_t)nc * pmx::kGeoStride, 0.0);
    for (int c = 0; c < nc; ++c) {
      const pmx::AffineMap a = pmx::affine_map(m, level, c);
      double* g = &geo[(size_t)c * pmx::kGeoStride];
      for (int r = 0; r < 3; ++r) g[r] = a.x0[r];
      for (int k = 0; k < 9; ++k) { g[3 + k] = a.J[k]; g[12 + k] = a.Jinv[k]; }
      g[21] = a.detJ;
      g[22] = pmx::cell_diameter(m, level, c);
    }
    s->d_geo = upload(geo.data(), geo.size(), st);
    s->d_vdofs = upload(s->vs.global_numbers.data(), s->vs.global_numbers.size(), st);
    s->d_pdofs = upload(s->ps.global_numbers.data(), s->ps.global_numbers.size(), st);
    std::vector<int> order;
    pmx::color_cells(s->vs, order, s->color_ptr);
    s->d_cells = upload(order.data(), order.size(), st);
    s->d_vxyz = upload(s->vs.dof_xyz.data(), s->vs.dof_xyz.size(), st);
    const pmx::QuadRule rule = pmx::assembly_rule(s->vs.el, 0);   // Gauss4 (degree 3 * 2)
    s->nq = rule.n;
    const std::vector<double> ref = nse_ref_table(rule, s->ps.el);
    s->d_ref = upload(ref.data(), ref.size(), st);
    s->d_Brp = upload(s->Bp.row_ptr.data(), s->Bp.row_ptr.size(), st);
    s->d_Bcol = upload(s->Bp.col.data(), s->Bp.col.size(), st);
    s->d_B1 = dmalloc<double>(s->Bp.col.size());
    s->d_B2 = dmalloc<double>(s->Bp.col.size());
    s->d_B3 = dmalloc<double>(s->Bp.col.size());
    s->d_BTrp = upload(s->BTp.row_ptr.data(), s->BTp.row_ptr.size(), st);
    s->d_BTcol = upload(s->BTp.col.data(), s->BTp.col.size(), st);
    s->d_BTsrc = upload(s->BTsrc.data(), s->BTsrc.size(), st);
    s->d_BT1 = dmalloc<double>(s->BTp.col.size());
    s->d_BT2 = dmalloc<double>(s->BTp.col.size());
    s->d_BT3 = dmalloc<double>(s->BTp.col.size());
    s->d_Crp = upload(s->Cp.row_ptr.data(), s->Cp.row_ptr.size(), st);
    s->d_Ccol = upload(s->Cp.col.data(), s->Cp.col.size(), st);
    s->d_Cval = dmalloc<double>(s->Cp.col.size());
    s->d_Ckind = upload(s->Ckind.data(), s->Ckind.size(), st);
    s->d_Cpos = upload(s->Cpos.data(), s->Cpos.size(), st);
    s->d_partial = dmalloc<double>((size_t)2 * ((nc + 255) / 256) + 8);
    HIP_CHECK(hipStreamSynchronize(st));
    *out = s.release();
  });
}

int pmx_nse_destroy(pmx_nse* s) {
  return guarded([&] { delete s; });
}

int pmx_nse_set_velocity_sell_rows(pmx_nse* s, int min_rows) {
  return guarded([&] {
    require(s && s->A, "NULL argument");
    require(min_rows >= 0, "min_rows must be >= 0");
    s->A->A.sell3_min_rows = min_rows;
  });
}

int pmx_nse_get_info(const pmx_nse* s, pmx_nse_info* out) {
  return guarded([&] {
    require(s && out, "NULL argument");
    out->n_cells = s->vs.n_cells;
    out->n_u = s->n_u;
    out->n_u_active = s->n_u_active;
    out->n_p = s->n_p;
    out->nnz_a = (long long)s->A->pattern->col.size();
    out->nnz_b = (long long)s->Bp.col.size();
    out->n_combined = s->n_comb;
    out->nnz_combined = (long long)s->Cp.col.size();
    out->pressure_space = s->pressure_space;
    out->n_p_cell = s->np_cell;
    out->enclosed = s->enclosed ? 1 : 0;
  });
}

int pmx_nse_assemble(pmx_ctx* ctx, pmx_nse* s, const pmx_nse_params* p, const double* dev_w, double* dev_rhs) {
  return guarded([&] {
    require(ctx && s && p, "NULL argument");
    require(s->ctx == ctx, "system was created on another context");
    require(p->example == s->example, "the system was created for another example (boundary conditions)");
    hipStream_t st = ctx->stream;
    const size_t na = s->A->pattern->col.size(), nb = s->Bp.col.size();
    pmx::vec_fill((int)na, 0.0, s->A->A.val, st);
    pmx::vec_fill((int)nb, 0.0, s->d_B1, st);
    pmx::vec_fill((int)nb, 0.0, s->d_B2, st);
    pmx::vec_fill((int)nb, 0.0, s->d_B3, st);
    if (dev_rhs) pmx::vec_fill(s->n_comb, 0.0, dev_rhs, st);
    s->A->sell_valid = false;
    pmx::NseAsm a;
    a.nq = s->nq;
    a.n_u = s->n_u;
    a.n_u_active = s->n_u_active;
    a.example = nse_component(s->example, 0);
    a.np = s->np_cell;
    a.nu = p->nu;
    a.convection = dev_w != nullptr;
    a.ref = s->d_ref;
    a.geo = s->d_geo;
    a.vdofs = s->d_vdofs;
    a.pdofs = s->d_pdofs;
    a.w = dev_w;
    a.A = s->A->A;
    a.Brp = s->d_Brp;
    a.Bcol = s->d_Bcol;
    a.B1 = s->d_B1;
    a.B2 = s->d_B2;
    a.B3 = s->d_B3;
    a.rhs = dev_rhs;
    for (size_t k = 0; k + 1 < s->color_ptr.size(); ++k)
      pmx::assemble_nse(a, s->d_cells + s->color_ptr[k], s->color_ptr[k + 1] - s->color_ptr[k], st);
    if (dev_rhs) pmx::nse_dirichlet_rhs(nse_component(s->example, 0), s->n_u, s->n_u_active, s->d_vxyz, dev_rhs, st);
    nse_refresh(s, st);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_nse_apply(pmx_ctx* ctx, const pmx_nse* s, const double* x, double* y) {
  return guarded([&] {
    require(ctx && s && x && y && x != y, "invalid argument");
    pmx::vec_fill(s->n_comb, 0.0, y, ctx->stream);
    nse_apply_scaled_add(s, x, y, 1.0, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_nse_apply_scaled_add(pmx_ctx* ctx, const pmx_nse* s, const double* x, double* y, double a) {
  return guarded([&] {
    require(ctx && s && x && y && x != y, "invalid argument");
    nse_apply_scaled_add(s, x, y, a, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_nse_get_structure(const pmx_nse* s, int which, int* nnz, int* host_row_ptr, int* host_col) {
  return guarded([&] {
    require(s && nnz, "NULL argument");
    const pmx::CSRPattern* P = (which == PMX_NSE_BLOCK_A) ? s->A->pattern.get()
                               : (which == PMX_NSE_BLOCK_B) ? &s->Bp
                               : (which == PMX_NSE_BLOCK_BT) ? &s->BTp
                               : (which == PMX_NSE_COMBINED) ? &s->Cp : nullptr;
    require(P != nullptr, "unknown block");
    *nnz = (int)P->col.size();
    if (host_row_ptr) std::memcpy(host_row_ptr, P->row_ptr.data(), P->row_ptr.size() * sizeof(int));
    if (host_col) std::memcpy(host_col, P->col.data(), P->col.size() * sizeof(int));
  });
}

int pmx_nse_get_values(pmx_ctx* ctx, const pmx_nse* s, int which, double* host_out) {
  return guarded([&] {
    require(ctx && s && host_out, "NULL argument");
    const double* src = nullptr;
    size_t n = 0;
    switch (which) {
      case PMX_NSE_BLOCK_A: src = s->A->A.val; n = s->A->pattern->col.size(); break;
      case PMX_NSE_BLOCK_B1: src = s->d_B1; n = s->Bp.col.size(); break;
      case PMX_NSE_BLOCK_B2: src = s->d_B2; n = s->Bp.col.size(); break;
      case PMX_NSE_BLOCK_B3: src = s->d_B3; n = s->Bp.col.size(); break;
      case PMX_NSE_COMBINED: src = s->d_Cval; n = s->Cp.col.size(); break;
      case PMX_NSE_BLOCK_B1T: src = s->d_BT1; n = s->BTp.col.size(); break;
      case PMX_NSE_BLOCK_B2T: src = s->d_BT2; n = s->BTp.col.size(); break;
      case PMX_NSE_BLOCK_B3T: src = s->d_BT3; n = s->BTp.col.size(); break;
      default: break;
    }
    require(src != nullptr, "unknown block");
    HIP_CHECK(hipMemcpyAsync(host_out, src, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (which == PMX_NSE_BLOCK_B1T || which == PMX_NSE_BLOCK_B2T || which == PMX_NSE_BLOCK_B3T) {
      // the B^T blocks act on the active velocity rows only: their Dirichlet
      // rows are empty, as the rectangular blocks of NSTYPE 4 are assembled
      for (int r = s->n_u_active; r < s->n_u; ++r)
        for (int q = s->BTp.row_ptr[r]; q < s->BTp.row_ptr[r + 1]; ++q) host_out[q] = 0.0;
    }
  });
}

int pmx_nse_get_global_numbers(const pmx_nse* s, int* host_velocity, int* host_pressure) {
  return guarded([&] {
    require(s, "NULL argument");
    if (host_velocity)
      std::memcpy(host_velocity, s->vs.global_numbers.data(), s->vs.global_numbers.size() * sizeof(int));
    if (host_pressure)
      std::memcpy(host_pressure, s->ps.global_numbers.data(), s->ps.global_numbers.size() * sizeof(int));
  });
}

int pmx_nse_interpolate(pmx_ctx* ctx, const pmx_nse* s, int example, double* dev_x) {
  return guarded([&] {
    require(ctx && s && dev_x, "NULL argument");
    require(example == PMX_EX_NSE_U2P1 || example == PMX_EX_NSE_U0P0, "unknown Navier-Stokes example");
    hipStream_t st = ctx->stream;
    for (int k = 0; k < 3; ++k)
      pmx::interpolate_example(nse_component(example, k), 0.0, s->n_u, s->d_vxyz, dev_x + (size_t)k * s->n_u, st);
    if (s->ps.el.disc) {
      // nodal functionals of every cell (NF_D_H_P1_3D_EvalAll) on the host
      std::vector<double> p((size_t)s->n_p);
      double pts[24];
      pmx::p1disc_nf_points(pts);
      for (int c = 0; c < s->ps.n_cells; ++c) {
        const pmx::AffineMap a = pmx::affine_map(s->dom->mesh, s->level, c);
        double v8[8], f[4];
        for (int i = 0; i < 8; ++i) {
          double x[3], e[6];
          pmx::map_to_phys(a, pts + 3 * i, x);
          pmx::example_exact(nse_component(example, 3), x[0], x[1], x[2], e);
          v8[i] = e[0];
        }
        pmx::p1disc_nf_apply(v8, f);
        for (int k = 0; k < 4; ++k) p[s->ps.global_numbers[(size_t)c * 4 + k]] = f[k];
      }
      HIP_CHECK(hipMemcpyAsync(dev_x + (size_t)3 * s->n_u, p.data(), p.size() * sizeof(double),
                               hipMemcpyHostToDevice, st));
      HIP_CHECK(hipStreamSynchronize(st));
    } else {
      double* pxyz = upload(s->ps.dof_xyz.data(), s->ps.dof_xyz.size(), st);
      pmx::interpolate_example(nse_component(example, 3), 0.0, s->n_p, pxyz, dev_x + (size_t)3 * s->n_u, st);
      HIP_CHECK(hipStreamSynchronize(st));
      dfree(pxyz);
    }
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_nse_restrict_velocity(pmx_ctx* ctx, pmx_nse* coarse, const pmx_nse* fine, const double* dev_fine_u,
                              double* dev_coarse_u) {
  return guarded([&] {
    require(ctx && coarse && fine && dev_fine_u && dev_coarse_u, "NULL argument");
    require(coarse->dom == fine->dom && fine->level == coarse->level + 1,
            "restriction needs consecutive levels of one domain");
    hipStream_t st = ctx->stream;
    auto it = coarse->restrict_from.find(fine);
    if (it == coarse->restrict_from.end()) {
      pmx::CSRPattern RF;
      std::vector<double> rv;
      pmx::build_function_restriction(coarse->vs, fine->vs, RF, rv);
      pmx::DevCSR D{};
      D.n = RF.n_rows;
      D.m = RF.n_cols;
      D.nnz = (int)RF.col.size();
      D.active = RF.n_rows;
      D.lpr = 8;
      D.row_ptr = upload(RF.row_ptr.data(), RF.row_ptr.size(), st);
      D.col = upload(RF.col.data(), RF.col.size(), st);
      D.val = upload(rv.data(), rv.size(), st);
      it = coarse->restrict_from.emplace(fine, D).first;
    }
    for (int k = 0; k < 3; ++k)
      pmx::csr_spmv(it->second, dev_fine_u + (size_t)k * fine->n_u, dev_coarse_u + (size_t)k * coarse->n_u, 1.0,
                    0.0, coarse->n_u, 0, st);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_nse_project_pressure(pmx_ctx* ctx, pmx_nse* s, double* dev_x) {
  return guarded([&] {
    require(ctx && s && dev_x, "NULL argument");
    pmx::pressure_project_l20(s->vs.n_cells, s->nq, s->np_cell, s->d_ref, s->d_geo, s->d_pdofs, s->n_p,
                              dev_x + (size_t)3 * s->n_u, s->d_partial, ctx->stream);
    HIP_CHECK(hipGetLastError());
  });
}

int pmx_nse_errors(pmx_ctx* ctx, pmx_nse* s, int example, const double* dev_x, double* host_err) {
  return guarded([&] {
    require(ctx && s && dev_x && host_err, "NULL argument");
    require(example == PMX_EX_NSE_U2P1 || example == PMX_EX_NSE_U0P0, "unknown Navier-Stokes example");
    hipStream_t st = ctx->stream;
    double l2u = 0, h1u = 0;
    for (int k = 0; k < 4; ++k) {
      const pmx::FESpace& fs = (k < 3) ? s->vs : s->ps;
      pmx::AssemblyData ad;
      ad.dim = 3;
      ad.ndof = fs.ndof_cell;
      ad.ref = scalar_ref(s, fs, &ad.nq);
      ad.n_cells = fs.n_cells;
      ad.geo = s->d_geo;
      ad.dofs = (k < 3) ? s->d_vdofs : s->d_pdofs;
      pmx::CDCoeffs co;
      co.example = nse_component(example, k);
      const int nb = (ad.n_cells + 255) / 256;
      double* partial = dmalloc<double>((size_t)nb * 4);
      pmx::cd_errors(ad, co, dev_x + (size_t)k * s->n_u, partial, ctx->scal, st);
      double e[4];
      HIP_CHECK(hipMemcpyAsync(e, ctx->scal, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      dfree(partial);
      if (k < 3) {
        l2u += e[0] * e[0];
        h1u += e[1] * e[1];
      } else {
        host_err[3] = e[0];
        host_err[4] = e[1];
      }
    }
    host_err[0] = std::sqrt(l2u);
    host_err[1] = std::sqrt(h1u);
    host_err[2] = 0.0;
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
