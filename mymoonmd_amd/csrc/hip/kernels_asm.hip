// Element-batched quadrature assembly of the convection-diffusion operator
// into the reference CSR pattern (gfx950).
//
// Restates, per cell: LocalAssembling<d>::GetLocalForms
// (src/AssembleRoutines/LocalAssembling.C:668-764) with the ConvDiff forms
// BilinearAssembleGalerkin<d> and BilinearAssemble_SD<d>
// (src/AssembleRoutines/ConvDiff.C:1794-2011) and the scatter of Assemble3D /
// Assemble2D (src/AssembleRoutines/Assemble3D.C:319-409, 467-507):
//   A_ij += w|J| (eps grad phi_j . grad phi_i + (b.grad phi_j) phi_i + c phi_j phi_i)
//   A_ij += w|J| delta (-eps lap phi_j + b.grad phi_j + c phi_j)(b.grad phi_i)  [SUPG]
//   f_i  += w|J| f phi_i (+ w|J| delta (b.grad phi_i) f)                         [SUPG]
// rows of Dirichlet DOFs get their diagonal set to 1.
//
// Layout: a 256-thread block holds 256/SLOT cells; the SLOT lanes of a cell
// own one local row (test function) each.  Per quadrature point the lanes
// stage the physical basis data of their cell in LDS, then every lane
// accumulates its 27-entry row in registers.  Cells of one colour share no
// DOF, so the scatter into the CSR row (binary search of the sorted columns)
// is race free without atomics.
#include <algorithm>
#include <mutex>
#include <stdexcept>

#include "common.hpp"
#include "examples.hpp"
#include "launch.hpp"

namespace pmx {

namespace {

template <int DIM>
struct Dims {
  static constexpr int NC = (DIM == 3) ? 6 : 3;   // second-derivative components
};

__device__ __forceinline__ int find_col(const int* __restrict__ ci, int s, int e, int c) {
  int lo = s, hi = e - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ci[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// per-(cell, q) coefficients: eps, b1, b2, b3, c, f, delta, mult
template <int DIM>
__device__ void coeffs_at(const CDCoeffs& co, const double* geo, const double* xi, double w,
                          int cell, int q, int nq, double* out) {
  double x[3] = {0, 0, 0};
  for (int r = 0; r < DIM; ++r) {
    double s = geo[r];
    for (int c = 0; c < DIM; ++c) s += geo[3 + r * DIM + c] * xi[c];
    x[r] = s;
  }
  double eps = co.eps, b[3] = {co.b[0], co.b[1], co.b[2]}, c = co.c, f;
  const double t = co.time_dev ? *co.time_dev : co.time;
  if (co.prob_coef) {
    const double* pc = co.prob_coef + 4 * co.cell_prob[cell];
    eps = pc[0]; b[0] = pc[1]; b[1] = pc[2]; b[2] = 0.0; c = pc[3];
  }
  if (co.table) {
    const double* t = co.table + ((size_t)cell * nq + q) * 8;
    eps = t[0]; b[0] = t[1]; b[1] = t[2]; b[2] = (DIM == 3) ? t[3] : 0.0; c = t[4]; f = t[5];
  } else {
    f = example_rhs(co.example, x[0], x[1], x[2], eps, b, c, co.f0, t);
  }
  const double detJ = geo[21], hK = geo[22];
  out[0] = eps; out[1] = b[0]; out[2] = b[1]; out[3] = b[2]; out[4] = c; out[5] = f;
  out[6] = co.supg ? sdfem_delta(hK, eps, b, DIM, c, co.delta0, co.poly_degree) : 0.0;
  out[7] = w * fabs(detJ);
}

template <int ND, int DIM, int SLOT>
__global__ __launch_bounds__(256) void k_assemble_cd(AssemblyData ad, CDCoeffs co,
                                                     const int* __restrict__ cells, int n,
                                                     DevCSR A, double* __restrict__ rhs) {
  constexpr int NC = Dims<DIM>::NC;
  constexpr int CPB = 256 / SLOT;
  constexpr int RS = 1 + DIM + ND * (1 + DIM + NC);   // ref table stride per q
  __shared__ double s_u[CPB][ND];
  __shared__ double s_g[CPB][DIM][ND];
  __shared__ double s_l[CPB][ND];
  __shared__ double s_c[CPB][8];
  const int slot = threadIdx.x / SLOT;
  const int i = threadIdx.x % SLOT;
  const int ci = blockIdx.x * CPB + slot;
  const bool valid = ci < n;
  const int cell = valid ? cells[ci] : 0;
  const double* geo = ad.geo + (size_t)cell * kGeoStride;
  double Jinv[DIM * DIM];
#pragma unroll
  for (int k = 0; k < DIM * DIM; ++k) Jinv[k] = geo[12 + k];
  // G = Jinv Jinv^T for the physical Laplacian on affine cells
  double G[DIM * DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k)
#pragma unroll
    for (int l = 0; l < DIM; ++l) {
      double s = 0;
#pragma unroll
      for (int a = 0; a < DIM; ++a) s += Jinv[k * DIM + a] * Jinv[l * DIM + a];
      G[k * DIM + l] = s;
    }
  double arow[ND];
#pragma unroll
  for (int j = 0; j < ND; ++j) arow[j] = 0.0;
  double frow = 0.0;

  for (int q = 0; q < ad.nq; ++q) {
    const double* rq = ad.ref + (size_t)q * RS;
    if (valid && i < ND) {
      const double* dphi = rq + 1 + DIM + ND;
      s_u[slot][i] = rq[1 + DIM + i];
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += dphi[k * ND + i] * Jinv[k * DIM + a];
        s_g[slot][a][i] = s;
      }
      if (co.supg) {
        const double* d2 = rq + 1 + DIM + ND * (1 + DIM);
        double lap;
        if (DIM == 3) {
          lap = d2[0 * ND + i] * G[0] + d2[3 * ND + i] * G[4] + d2[5 * ND + i] * G[8] +
                2.0 * (d2[1 * ND + i] * G[1] + d2[2 * ND + i] * G[2] + d2[4 * ND + i] * G[5]);
        } else {
          lap = d2[0 * ND + i] * G[0] + d2[2 * ND + i] * G[3] + 2.0 * d2[1 * ND + i] * G[1];
        }
        s_l[slot][i] = lap;
      }
    }
    if (valid && i == 0) coeffs_at<DIM>(co, geo, rq + 1, rq[0], cell, q, ad.nq, s_c[slot]);
    __syncthreads();
    if (valid && i < ND) {
      const double eps = s_c[slot][0], b1 = s_c[slot][1], b2 = s_c[slot][2];
      const double b3 = (DIM == 3) ? s_c[slot][3] : 0.0;
      const double c = s_c[slot][4], f = s_c[slot][5], delta = s_c[slot][6], mult = s_c[slot][7];
      const double ti = s_u[slot][i];
      const double tx = s_g[slot][0][i], ty = s_g[slot][1][i];
      const double tz = (DIM == 3) ? s_g[slot][DIM - 1][i] : 0.0;
      const double bgradv = b1 * tx + b2 * ty + b3 * tz;
      // GLS tests with L phi_i = -eps lap phi_i + b.grad phi_i + c phi_i
      double Lu = 0.0;
      if (co.supg == 2) {
        if (DIM == 3) {
          Lu = -eps * s_l[slot][i];
          Lu += b1 * tx + b2 * ty + b3 * tz + c * ti;
        } else {
          Lu = -eps * s_l[slot][i] + b1 * tx + b2 * ty + c * ti;
        }
      }
      frow += mult * ti * f;
      if (co.supg == 1) frow += mult * (delta * bgradv) * f;
      else if (co.supg == 2) frow += mult * delta * Lu * f;
      const double gls_w = mult * delta * Lu;   // 3D factor of BilinearAssemble_GLS<3>
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        const double ax = s_g[slot][0][j], ay = s_g[slot][1][j];
        const double az = (DIM == 3) ? s_g[slot][DIM - 1][j] : 0.0;
        const double aj = s_u[slot][j];
        const double bgu = b1 * ax + b2 * ay + b3 * az;
        double val = eps * (tx * ax + ty * ay + tz * az);
        val += bgu * ti;
        val += c * aj * ti;
        arow[j] += val * mult;
        if (co.supg == 1) {
          double v2 = -eps * s_l[slot][j];
          v2 += bgu + c * aj;
          arow[j] += v2 * delta * bgradv * mult;
        } else if (co.supg == 2) {
          if (DIM == 3) {
            double v2 = -eps * s_l[slot][j];
            v2 += bgu + c * aj;
            arow[j] += v2 * gls_w;
          } else {
            const double v2 = delta * (-eps * s_l[slot][j] + b1 * ax + b2 * ay + c * aj) * Lu;
            arow[j] += v2 * mult;
          }
        }
      }
    }
    __syncthreads();
  }
  if (!valid || i >= ND) return;
  const int* dofs = ad.dofs + (size_t)cell * ND;
  const int r = dofs[i];
  const int rs = A.row_ptr[r], re = A.row_ptr[r + 1];
  if (r < A.active) {
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      const int pos = find_col(A.col, rs, re, dofs[j]);
      A.val[pos] += arow[j];
    }
    rhs[r] += frow;
  } else {
    A.val[find_col(A.col, rs, re, r)] = 1.0;
  }
}

template <int ND, int DIM, int SLOT>
void launch_asm(const AssemblyData& ad, const CDCoeffs& co, const int* cells, int n,
                const DevCSR& A, double* rhs, hipStream_t st) {
  constexpr int CPB = 256 / SLOT;
  const int grid = (n + CPB - 1) / CPB;
  if (grid > 0)
    hipLaunchKernelGGL((k_assemble_cd<ND, DIM, SLOT>), dim3(grid), dim3(256), 0, st, ad, co, cells, n,
                       A, rhs);
}

__global__ void k_dirichlet_example(int ex, double t, int n, const double* __restrict__ xyz,
                                    double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = example_dirichlet(ex, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], t);
}

__global__ void k_interpolate_example(int ex, double t, int n, const double* __restrict__ xyz,
                                      double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    double v[6];
    example_exact(ex, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], v, t);
    out[k] = v[0];
  }
}

// Time-dependent CD forms (src/AssembleRoutines/CD_local_assembling_routines.C):
//   TCDStiff   A_ij += w|J| (eps grad phi_j . grad phi_i + (b.grad phi_j) phi_i + c phi_j phi_i)
//   TCDMass    M_ij += w|J| phi_i phi_j
//   TCDRhs     f_i  += w|J| f phi_i
// and in 2D with SUPG, with e = eps t1 tau, B = b t1 tau, C = 1 + c t1 tau,
// delta = Compute_SDFEM_delta(hK, e, B):
//   TCDStiffSUPG  A_ij += w|J| delta (B.grad phi_i) (B.grad phi_j + C phi_j - e lap phi_j)
//   TCDMassSUPG   M_ij += w|J| delta t1 tau (B.grad phi_i) phi_j
//   TCDRhsSUPG    f_i  += w|J| delta (B.grad phi_i) f
// Same lane layout as k_assemble_cd; each lane keeps its stiffness row, mass
// row and load entry.
template <int ND, int DIM, int SLOT>
__global__ __launch_bounds__(256) void k_assemble_tcd(AssemblyData ad, CDCoeffs co,
                                                      const int* __restrict__ cells, int n,
                                                      int active, DevCSR A, DevCSR M,
                                                      double* __restrict__ rhs) {
  constexpr int NC = Dims<DIM>::NC;
  constexpr int CPB = 256 / SLOT;
  constexpr int RS = 1 + DIM + ND * (1 + DIM + NC);
  __shared__ double s_u[CPB][ND];
  __shared__ double s_g[CPB][DIM][ND];
  __shared__ double s_l[CPB][ND];
  __shared__ double s_c[CPB][12];
  const bool supg = co.supg && DIM == 2;
  const int slot = threadIdx.x / SLOT;
  const int i = threadIdx.x % SLOT;
  const int ci = blockIdx.x * CPB + slot;
  const bool valid = ci < n;
  const int cell = valid ? cells[ci] : 0;
  const double* geo = ad.geo + (size_t)cell * kGeoStride;
  double Jinv[DIM * DIM];
#pragma unroll
  for (int k = 0; k < DIM * DIM; ++k) Jinv[k] = geo[12 + k];
  double G[DIM * DIM];
#pragma unroll
  for (int k = 0; k < DIM; ++k)
#pragma unroll
    for (int l = 0; l < DIM; ++l) {
      double s = 0;
#pragma unroll
      for (int a = 0; a < DIM; ++a) s += Jinv[k * DIM + a] * Jinv[l * DIM + a];
      G[k * DIM + l] = s;
    }
  double arow[ND], mrow[ND];
#pragma unroll
  for (int j = 0; j < ND; ++j) { arow[j] = 0.0; mrow[j] = 0.0; }
  double frow = 0.0;

  for (int q = 0; q < ad.nq; ++q) {
    const double* rq = ad.ref + (size_t)q * RS;
    if (valid && i < ND) {
      const double* dphi = rq + 1 + DIM + ND;
      s_u[slot][i] = rq[1 + DIM + i];
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += dphi[k * ND + i] * Jinv[k * DIM + a];
        s_g[slot][a][i] = s;
      }
      if (supg) {
        const double* d2 = rq + 1 + DIM + ND * (1 + DIM);
        s_l[slot][i] = d2[0 * ND + i] * G[0] + d2[2 * ND + i] * G[DIM * DIM - 1] +
                       2.0 * d2[1 * ND + i] * G[1];
      }
    }
    if (valid && i == 0) {
      double cf[8];
      coeffs_at<DIM>(co, geo, rq + 1, rq[0], cell, q, ad.nq, cf);
      double* sc = s_c[slot];
      for (int k = 0; k < 6; ++k) sc[k] = cf[k];
      sc[6] = cf[7];   // w |det J|
      if (supg) {
        const double t1 = co.theta1, tau = co.tau;
        const double se = cf[0] * t1 * tau;
        const double sb[3] = {cf[1] * t1 * tau, cf[2] * t1 * tau, 0.0};
        const double scc = 1. + cf[4] * t1 * tau;
        const double hK = geo[22];
        const double delta = sdfem_delta(hK, se, sb, 2, scc, co.delta0, co.poly_degree);
        sc[7] = se; sc[8] = sb[0]; sc[9] = sb[1]; sc[10] = scc; sc[11] = delta;
      }
    }
    __syncthreads();
    if (valid && i < ND) {
      const double* sc = s_c[slot];
      const double eps = sc[0], b1 = sc[1], b2 = sc[2];
      const double b3 = (DIM == 3) ? sc[3] : 0.0;
      const double c = sc[4], f = sc[5], mult = sc[6];
      const double ti = s_u[slot][i];
      const double tx = s_g[slot][0][i], ty = s_g[slot][1][i];
      const double tz = (DIM == 3) ? s_g[slot][DIM - 1][i] : 0.0;
      frow += mult * ti * f;
#pragma unroll
      for (int j = 0; j < ND; ++j) {
        const double ax = s_g[slot][0][j], ay = s_g[slot][1][j];
        const double az = (DIM == 3) ? s_g[slot][DIM - 1][j] : 0.0;
        const double aj = s_u[slot][j];
        double val = eps * (tx * ax + ty * ay + tz * az);
        val += (b1 * ax + b2 * ay + b3 * az) * ti;
        val += c * ti * aj;
        arow[j] += mult * val;
        mrow[j] += mult * ti * aj;
      }
      if (supg) {
        const double se = sc[7], sb1 = sc[8], sb2 = sc[9], scc = sc[10], delta = sc[11];
        const double bgradv = delta * (sb1 * tx + sb2 * ty);
#pragma unroll
        for (int j = 0; j < ND; ++j) {
          const double ax = s_g[slot][0][j], ay = s_g[slot][1][j];
          double val = (sb1 * ax + sb2 * ay) * bgradv;
          val += scc * bgradv * s_u[slot][j];
          val -= se * s_l[slot][j] * bgradv;
          arow[j] += mult * val;
        }
        const double bgm = (delta * co.theta1 * co.tau) * (sb1 * tx + sb2 * ty);
#pragma unroll
        for (int j = 0; j < ND; ++j) mrow[j] += mult * s_u[slot][j] * bgm;
        frow += mult * f * bgradv;
      }
    }
    __syncthreads();
  }
  if (!valid || i >= ND) return;
  const int* dofs = ad.dofs + (size_t)cell * ND;
  const int r = dofs[i];
  const int act_end = co.prob_active_end ? co.prob_active_end[co.cell_prob[cell]] : active;
  if (r < act_end) {
    if (A.val) {
      const int rs = A.row_ptr[r], re = A.row_ptr[r + 1];
#pragma unroll
      for (int j = 0; j < ND; ++j) A.val[find_col(A.col, rs, re, dofs[j])] += arow[j];
    }
    if (M.val) {
      const int rs = M.row_ptr[r], re = M.row_ptr[r + 1];
#pragma unroll
      for (int j = 0; j < ND; ++j) M.val[find_col(M.col, rs, re, dofs[j])] += mrow[j];
    }
    if (rhs) rhs[r] += frow;
  } else {
    if (A.val) A.val[find_col(A.col, A.row_ptr[r], A.row_ptr[r + 1], r)] = 1.0;
    if (M.val) M.val[find_col(M.col, M.row_ptr[r], M.row_ptr[r + 1], r)] = 1.0;
  }
}

template <int ND, int DIM, int SLOT>
void launch_tcd(const AssemblyData& ad, const CDCoeffs& co, const int* cells, int n, int active,
                const DevCSR& A, const DevCSR& M, double* rhs, hipStream_t st) {
  constexpr int CPB = 256 / SLOT;
  const int grid = (n + CPB - 1) / CPB;
  if (grid > 0)
    hipLaunchKernelGGL((k_assemble_tcd<ND, DIM, SLOT>), dim3(grid), dim3(256), 0, st, ad, co, cells, n,
                       active, A, M, rhs);
}

// one thread per cell; partial sums per block
template <int ND, int DIM>
__global__ __launch_bounds__(256) void k_cd_errors(AssemblyData ad, CDCoeffs co,
                                                   const double* __restrict__ u,
                                                   double* __restrict__ partial) {
  constexpr int NC = Dims<DIM>::NC;
  constexpr int RS = 1 + DIM + ND * (1 + DIM + NC);
  __shared__ double red[16];
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  double e[4] = {0, 0, 0, 0};
  if (cell < ad.n_cells) {
    const double* geo = ad.geo + (size_t)cell * kGeoStride;
    double uloc[ND];
    for (int j = 0; j < ND; ++j) uloc[j] = u[ad.dofs[(size_t)cell * ND + j]];
    for (int q = 0; q < ad.nq; ++q) {
      const double* rq = ad.ref + (size_t)q * RS;
      double cf[8];
      coeffs_at<DIM>(co, geo, rq + 1, rq[0], cell, q, ad.nq, cf);
      double val = 0, g[3] = {0, 0, 0};
      const double* dphi = rq + 1 + DIM + ND;
      for (int j = 0; j < ND; ++j) {
        val += uloc[j] * rq[1 + DIM + j];
        for (int a = 0; a < DIM; ++a) {
          double s = 0;
          for (int k = 0; k < DIM; ++k) s += dphi[k * ND + j] * geo[12 + k * DIM + a];
          g[a] += uloc[j] * s;
        }
      }
      double x[3] = {0, 0, 0};
      for (int r = 0; r < DIM; ++r) {
        double s = geo[r];
        for (int c = 0; c < DIM; ++c) s += geo[3 + r * DIM + c] * rq[1 + c];
        x[r] = s;
      }
      double ex[6];
      example_exact(co.example, x[0], x[1], x[2], ex, co.time);
      const double w = cf[7];
      const double e0 = val - ex[0], e1 = g[0] - ex[1], e2 = g[1] - ex[2];
      const double e3 = (DIM == 3) ? g[2] - ex[3] : 0.0;
      e[0] += w * e0 * e0;
      e[1] += w * (e1 * e1 + e2 * e2 + e3 * e3);
      const double e4 = cf[1] * e1 + cf[2] * e2 + cf[3] * e3;
      e[2] += w * (cf[0] * (e1 * e1 + e2 * e2 + e3 * e3) + cf[4] * e0 * e0 + cf[6] * e4 * e4);
      e[3] = fmax(e[3], fabs(e0));
    }
  }
  for (int k = 0; k < 3; ++k) {
    const double s = block_sum(e[k], red);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * 4 + k] = s;
  }
  // max via the same reduction tree on squared-free values
  double m = e[3];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) mm = fmax(mm, red[w]);
    partial[(size_t)blockIdx.x * 4 + 3] = mm;
  }
}

__global__ __launch_bounds__(1024) void k_reduce_errors(int nb, const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double red[16];
  double s[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    for (int k = 0; k < 3; ++k) s[k] += partial[(size_t)b * 4 + k];
    s[3] = fmax(s[3], partial[(size_t)b * 4 + 3]);
  }
  for (int k = 0; k < 3; ++k) {
    const double t = block_sum(s[k], red);
    if (threadIdx.x == 0) out[k] = sqrt(t);
  }
  double m = s[3];
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) mm = fmax(mm, red[w]);
    out[3] = mm;
  }
}

}  // namespace

// Q2 hexahedra on the matrix cores: with per quadrature point q the staged
// physical basis data g = grad phi, the Galerkin/SUPG/GLS element matrix is
//   A_ij = sum_q m_q [ eps g_i.g_j + (phi_i + delta t_i)(b.g_j + c phi_j)
//                      + delta t_i (-eps lap phi_j) ]
// (t = b.grad phi for SUPG, L phi = -eps lap phi + b.grad phi + c phi for
// GLS, delta = 0 for Galerkin): a 27 x K times K x 27 product with K = 5 per
// quadrature point (three gradient channels, the convection-reaction channel,
// the streamline-diffusion channel).  One wavefront per cell accumulates it as
// 2 x 2 tiles of v_mfma_f64_16x16x4f64 (rows/columns 27..31 zero), FP64
// throughout; the load vector is f_i = sum_q m_q f (phi_i + delta t_i).
// Quadrature points are staged 4 at a time (5 MFMA k-steps per chunk).
// Lane layout of the f64 16x16x4 MFMA: A operand lane l = A[l % 16][l / 16],
// B operand lane l = B[l / 16][l % 16], result element t of lane l =
// D[4 t + l / 16][l % 16].
typedef double v4d_t __attribute__((ext_vector_type(4)));


template <int QMAX>
__global__ __launch_bounds__(256) void k_assemble_cd_q2_mfma(AssemblyData ad, CDCoeffs co,
                                                             const int* __restrict__ cells, int n, DevCSR A,
                                                             double* __restrict__ rhs) {
  constexpr int ND = 27, DIM = 3, NC = 6, QC = 4, NV = 7;
  constexpr int RS = 1 + DIM + ND * (1 + DIM + NC);
  __shared__ double s_p[4][QC][ND][NV];   // per wave: g1 g2 g3, phi + delta t, delta t, b.g + c phi, -eps lap
  __shared__ double s_c[4][QMAX][8];      // per wave: eps b1 b2 b3 c f delta mult of every quadrature point
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = blockIdx.x * 4 + wv;
  const bool valid = ci < n;
  const int cell = valid ? cells[ci] : cells[0];   // a real cell for the reads; no writes when !valid
  const double* geo = ad.geo + (size_t)cell * kGeoStride;
  double Jinv[9], G[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Jinv[k] = geo[12 + k];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      double t = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a) t += Jinv[k * 3 + a] * Jinv[l * 3 + a];
      G[k * 3 + l] = t;
    }
  if (lane < ad.nq) {
    const double* rq = ad.ref + (size_t)lane * RS;
    coeffs_at<3>(co, geo, rq + 1, rq[0], cell, lane, ad.nq, s_c[wv][lane]);
  }
  wave_sync();
  v4d_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d_t{0.0, 0.0, 0.0, 0.0};
  double frow = 0.0;
  for (int q0 = 0; q0 < ad.nq; q0 += QC) {
    for (int e = lane; e < QC * ND; e += 64) {
      const int qq = e / ND, i = e % ND, q = q0 + qq;
      double* out = s_p[wv][qq][i];
      if (q < ad.nq) {
        const double* rq = ad.ref + (size_t)q * RS;
        const double* dphi = rq + 1 + DIM + ND;
        const double* d2 = rq + 1 + DIM + ND * (1 + DIM);
        const double phi = rq[1 + DIM + i];
        double g[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          double t = 0;
#pragma unroll
          for (int k = 0; k < 3; ++k) t += dphi[k * ND + i] * Jinv[k * 3 + a];
          g[a] = t;
        }
        const double lap = d2[0 * ND + i] * G[0] + d2[3 * ND + i] * G[4] + d2[5 * ND + i] * G[8] +
                           2.0 * (d2[1 * ND + i] * G[1] + d2[2 * ND + i] * G[2] + d2[4 * ND + i] * G[5]);
        const double* c = s_c[wv][q];
        const double eps = c[0], cc = c[4], delta = c[6];
        const double bg = c[1] * g[0] + c[2] * g[1] + c[3] * g[2];
        double t = 0.0;
        if (co.supg == 1) t = bg;
        else if (co.supg == 2) { t = -eps * lap; t += bg + cc * phi; }
        out[0] = g[0]; out[1] = g[1]; out[2] = g[2];
        out[3] = phi + delta * t;
        out[4] = delta * t;
        out[5] = bg + cc * phi;
        out[6] = -eps * lap;
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = 0.0;
      }
    }
    wave_sync();
    if (lane < ND)
      for (int qq = 0; qq < QC && q0 + qq < ad.nq; ++qq) {
        const double* c = s_c[wv][q0 + qq];
        frow += c[7] * c[5] * s_p[wv][qq][lane][3];
      }
    const int r16 = lane & 15;
#pragma unroll
    for (int st = 0; st < 5 * QC / 4; ++st) {
      const int k = 4 * st + (lane >> 4);
      const int qq = k / 5, ch = k % 5, q = q0 + qq;
      double m = 0.0, me = 0.0;
      if (q < ad.nq) {
        m = s_c[wv][q][7];
        me = m * s_c[wv][q][0];
      }
      const int ta = (ch < 3) ? ch : ch;       // test channel value: g_ch, phi + delta t, delta t
      const int tb = (ch < 3) ? ch : ch + 2;   // trial channel value: g_ch, b.g + c phi, -eps lap
      const double sa = (ch < 3) ? me : m;
      double av[2], bv[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = 16 * h + r16;
        av[h] = (i < ND) ? sa * s_p[wv][qq][i][ta] : 0.0;
        bv[h] = (i < ND) ? s_p[wv][qq][i][tb] : 0.0;
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    wave_sync();
  }
  if (!valid) return;
  const int* dofs = ad.dofs + (size_t)cell * ND;
  if (ad.pos) {
    // precomputed positions: all loads (16 coalesced position loads, the 8
    // row DOFs, the 16 old values) are issued before any store, so the
    // read-modify-writes overlap instead of paying one memory latency each
    const int* pc = ad.pos + (size_t)cell * kQ2PosPerCell + lane;
    int p[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) p[k] = pc[k * 64];
    bool act[8];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = 16 * a + 4 * t + (lane >> 4);
        act[a * 4 + t] = dofs[i < ND ? i : 0] < A.active;
      }
    double old[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) old[k] = A.val[p[k] >= 0 ? p[k] : 0];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int k = (a * 2 + b) * 4 + t;
          // Dirichlet rows keep only their unit diagonal (the one slot stored)
          if (p[k] >= 0) A.val[p[k]] = act[a * 4 + t] ? old[k] + acc[a][b][t] : 1.0;
        }
  } else {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int i = 16 * a + 4 * t + (lane >> 4), j = 16 * b + (lane & 15);
          if (i >= ND || j >= ND) continue;
          const int r = dofs[i];
          const int rs = A.row_ptr[r], re = A.row_ptr[r + 1];
          if (r < A.active) A.val[find_col(A.col, rs, re, dofs[j])] += acc[a][b][t];
          else if (i == j) A.val[find_col(A.col, rs, re, r)] = 1.0;
        }
  }
  if (lane < ND) {
    const int r = dofs[lane];
    if (r < A.active) rhs[r] += frow;
  }
}

// position table of the Q2 matrix-core assembly: slot ((a * 2 + b) * 4 + t) * 64
// + lane of a cell holds the entry (i, j) = (16 a + 4 t + lane / 16,
// 16 b + lane % 16) of its element matrix
__global__ void k_q2_scatter_positions(int n_cells, const int* __restrict__ dofs, DevCSR A, int* __restrict__ pos) {
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (size_t)n_cells * kQ2PosPerCell) return;
  const int cell = (int)(g / kQ2PosPerCell), slot = (int)(g % kQ2PosPerCell);
  const int lane = slot & 63, tile = slot >> 6;   // tile = (a * 2 + b) * 4 + t
  const int t = tile & 3, b = (tile >> 2) & 1, a = tile >> 3;
  const int i = 16 * a + 4 * t + (lane >> 4), j = 16 * b + (lane & 15);
  int p = -1;
  if (i < 27 && j < 27) {
    const int r = dofs[(size_t)cell * 27 + i], c = dofs[(size_t)cell * 27 + j];
    if (r < A.active || r == c) p = find_col(A.col, A.row_ptr[r], A.row_ptr[r + 1], c);
  }
  pos[g] = p;
}

void q2_scatter_positions(const AssemblyData& ad, const DevCSR& A, int* pos, hipStream_t st) {
  const size_t total = (size_t)ad.n_cells * kQ2PosPerCell;
  if (total > 0)
    hipLaunchKernelGGL(k_q2_scatter_positions, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, ad.n_cells,
                       ad.dofs, A, pos);
}

// ---- Q2 hexahedra, coefficients constant on each cell: reference matrices
//
// On an affine cell with constant eps, b, c (the built-in examples; a
// per-point coefficient table goes to the quadrature kernel above) the
// Galerkin/SUPG element matrix of k_assemble_cd_q2_mfma expands, term by
// term, into
//   A_e = sum_r alpha_r(e) R_r,   R_r[i][j] = sum_q w_q u_r(q, i) v_r(q, j),
// with the same quadrature rule on the reference cell.  With beta = J^-1 b
// (b.grad phi = sum_k beta_k d_k phi), G = J^-1 J^-T, m = |det J|, D_k the
// reference derivatives and D_s, s = xx xy xz yy yz zz, the second ones:
//   r 0-2    D_k phi_i D_k phi_j                 m (eps G_kk + delta beta_k^2)
//   r 3-5    D_k phi_i D_l phi_j + (k <-> l)     m (eps G_kl + delta beta_k beta_l)   kl = 01 02 12
//   r 6-8    phi_i D_l phi_j                     m beta_l
//   r 9      phi_i phi_j                         m c
//   r 10-12  D_k phi_i phi_j                     m delta c beta_k
//   r 13-30  D_k phi_i D_s phi_j  (13 + 6k + s)  -m delta eps beta_k G'_s
// (G'_s = G_mm for s = mm, 2 G_mn for s = mn, m != n: -eps lap phi =
// -eps sum_s G'_s D_s phi; delta = 0 for Galerkin).  So all element matrices
// are one GEMM E[cells][27 x 32] = alpha[cells][32] R[32][27 x 32] (element
// rows padded to 32 entries), run on the matrix cores in FP64 or FP32 (FP32
// accumulation).  Element row i of cell e is written to the slot of the
// incidence (row dofs[e][i], cell e) in the order the rows read them, and a
// second kernel assembles every active CSR row once from its contiguous,
// aligned slots, adding them in increasing cell order (the order of
// Assemble3D's cell loop, Assemble3D.C:319-409) in LDS and writing the row
// with coalesced stores -- instead of a scattered read-modify-write per
// element entry in eight colour passes.

void q2_reference_matrices(const double* ref, int nq, double* R) {
  constexpr int ND = 27, RS = 1 + 3 + ND * 10;
  for (size_t k = 0; k < (size_t)kQ2Terms * kQ2N; ++k) R[k] = 0.0;
  static const int kl[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int q = 0; q < nq; ++q) {
    const double* rq = ref + (size_t)q * RS;
    const double w = rq[0];
    const double* phi = rq + 4;
    const double* d1 = rq + 4 + ND;            // d1[k * 27 + i]
    const double* d2 = rq + 4 + ND * 4;        // d2[s * 27 + i]
    for (int i = 0; i < ND; ++i)
      for (int j = 0; j < ND; ++j) {
        const int n = i * 32 + j;
        auto add = [&](int r, double v) { R[(size_t)r * kQ2N + n] += w * v; };
        for (int k = 0; k < 3; ++k) add(k, d1[k * ND + i] * d1[k * ND + j]);
        for (int t = 0; t < 3; ++t) {
          const int k = kl[t][0], l = kl[t][1];
          add(3 + t, d1[k * ND + i] * d1[l * ND + j] + d1[l * ND + i] * d1[k * ND + j]);
        }
        for (int l = 0; l < 3; ++l) add(6 + l, phi[i] * d1[l * ND + j]);
        add(9, phi[i] * phi[j]);
        for (int k = 0; k < 3; ++k) add(10 + k, d1[k * ND + i] * phi[j]);
        for (int k = 0; k < 3; ++k)
          for (int s = 0; s < 6; ++s) add(13 + 6 * k + s, d1[k * ND + i] * d2[s * ND + j]);
      }
  }
}

namespace {

// alpha_r of one cell (the table above) and (m, delta, beta) for its load vector
__device__ void q2_cell_terms(const CDCoeffs& co, const double* geo, double* al, double* cd) {
  double Ji[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Ji[k] = geo[12 + k];
  const double m = fabs(geo[21]), hK = geo[22];
  const double eps = co.eps, c = co.c;
  double beta[3], G[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) beta[k] = Ji[3 * k] * co.b[0] + Ji[3 * k + 1] * co.b[1] + Ji[3 * k + 2] * co.b[2];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int l = 0; l < 3; ++l)
      G[3 * k + l] = Ji[3 * k] * Ji[3 * l] + Ji[3 * k + 1] * Ji[3 * l + 1] + Ji[3 * k + 2] * Ji[3 * l + 2];
  const double delta = co.supg ? sdfem_delta(hK, eps, co.b, 3, c, co.delta0, co.poly_degree) : 0.0;
  al[0] = m * (eps * G[0] + delta * beta[0] * beta[0]);
  al[1] = m * (eps * G[4] + delta * beta[1] * beta[1]);
  al[2] = m * (eps * G[8] + delta * beta[2] * beta[2]);
  al[3] = m * (eps * G[1] + delta * beta[0] * beta[1]);
  al[4] = m * (eps * G[2] + delta * beta[0] * beta[2]);
  al[5] = m * (eps * G[5] + delta * beta[1] * beta[2]);
#pragma unroll
  for (int l = 0; l < 3; ++l) al[6 + l] = m * beta[l];
  al[9] = m * c;
#pragma unroll
  for (int k = 0; k < 3; ++k) al[10 + k] = m * delta * c * beta[k];
  const double Gs[6] = {G[0], 2.0 * G[1], 2.0 * G[2], G[4], 2.0 * G[5], G[8]};
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int s = 0; s < 6; ++s) al[13 + 6 * k + s] = -m * delta * eps * beta[k] * Gs[s];
  al[31] = 0.0;
  cd[0] = m;
  cd[1] = delta;
  cd[2] = beta[0];
  cd[3] = beta[1];
  cd[4] = beta[2];
}

typedef float v4f_t __attribute__((ext_vector_type(4)));

// one MFMA k-step of a 16 x 16 tile and the row of result element t of a lane
__device__ __forceinline__ v4d_t mfma16(double a, double b, v4d_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f_t mfma16(float a, float b, v4f_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// f64 16x16x4: element t of lane l is D[4 t + l / 16][l % 16];
// f32 16x16x4: element t of lane l is D[4 (l / 16) + t][l % 16]
__device__ __forceinline__ int mfma16_row(double, int t, int lane) { return 4 * t + (lane >> 4); }
__device__ __forceinline__ int mfma16_row(float, int t, int lane) { return 4 * (lane >> 4) + t; }
template <typename T> struct Acc16;
template <> struct Acc16<double> { typedef v4d_t type; };
template <> struct Acc16<float> { typedef v4f_t type; };

constexpr int kQ2GemmCells = 64;   // cells per workgroup of the element GEMM
constexpr int kQ2MaxQ = 64;        // quadrature points the load-vector stage holds

// 64 cells per 256-thread workgroup: (1) alpha and the load-vector data of
// each cell, (2) w_q f(x_q) at every quadrature point, (3) the element load
// vectors, (4) E = alpha R: every wave owns the N tiles nt = wave (mod 4) of
// all four 16-cell M tiles, with the four tiles' alpha operands in registers;
// element row i of cell e goes to its incidence slot (lane e of the wave
// reads the slot of cell e, the lanes holding a result fetch it).
template <typename T>
__global__ __launch_bounds__(256) void k_q2_element_gemm(AssemblyData ad, CDCoeffs co, Q2Gather g,
                                                         const T* __restrict__ R, T* __restrict__ E,
                                                         double* __restrict__ fe) {
  constexpr int ND = 27, RS = 1 + 3 + ND * 10, NC = kQ2GemmCells;
  __shared__ double s_al[NC][kQ2Terms + 1];
  __shared__ double s_cd[NC][6];
  // w_q f(x_q) per cell and point; after the load vectors the cells' slots
  __shared__ double s_wf[NC][kQ2MaxQ + 1];
  int* s_slot = reinterpret_cast<int*>(&s_wf[0][0]);
  const int tid = threadIdx.x, c0 = blockIdx.x * NC, nq = ad.nq, n_cells = ad.n_cells;
  if (tid < NC) {
    const int e = c0 + tid;
    if (e < n_cells) {
      q2_cell_terms(co, ad.geo + (size_t)e * kGeoStride, s_al[tid], s_cd[tid]);
    } else {
      for (int r = 0; r < kQ2Terms; ++r) s_al[tid][r] = 0.0;
      for (int r = 0; r < 6; ++r) s_cd[tid][r] = 0.0;
    }
  }
  const double t = co.time_dev ? *co.time_dev : co.time;
  for (int k = tid; k < NC * nq; k += 256) {
    const int lc = k / nq, q = k - lc * nq, e = c0 + lc;
    double v = 0.0;
    if (e < n_cells) {
      const double* geo = ad.geo + (size_t)e * kGeoStride;
      const double* rq = ad.ref + (size_t)q * RS;
      double x[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) x[r] = geo[r] + geo[3 + 3 * r] * rq[1] + geo[4 + 3 * r] * rq[2] + geo[5 + 3 * r] * rq[3];
      v = rq[0] * example_rhs(co.example, x[0], x[1], x[2], co.eps, co.b, co.c, co.f0, t);
    }
    s_wf[lc][q] = v;
  }
  __syncthreads();
  // fe = m sum_q w_q f_q (phi_i + delta beta.D phi_i) of every active (e, i):
  // thread t < 243 keeps the basis function i = t % 27 and sums the cells
  // t / 27 + 9 c, so each reference value it loads serves up to 8 cells
  if (tid < 9 * ND) {
    constexpr int CPT = (NC + 8) / 9;
    const int i = tid % ND, c1 = tid / ND;
    double s0[CPT], s1[CPT], s2[CPT], s3[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) s0[c] = s1[c] = s2[c] = s3[c] = 0.0;
    for (int q = 0; q < nq; ++q) {
      const double* rq = ad.ref + (size_t)q * RS + 4;
      const double p0 = rq[i], p1 = rq[ND + i], p2 = rq[2 * ND + i], p3 = rq[3 * ND + i];
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        const int lc = c1 + 9 * c;
        if (lc < NC) {
          const double wf = s_wf[lc][q];
          s0[c] += wf * p0;
          s1[c] += wf * p1;
          s2[c] += wf * p2;
          s3[c] += wf * p3;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int lc = c1 + 9 * c, e = c0 + lc;
      if (lc >= NC || e >= n_cells) continue;
      const int sl = g.slot[(size_t)e * ND + i];
      if (sl < 0) continue;
      const double* cd = s_cd[lc];
      fe[sl] = cd[0] * (s0[c] + cd[1] * (cd[2] * s1[c] + cd[3] * s2[c] + cd[4] * s3[c]));
    }
  }
  __syncthreads();
  for (int k = tid; k < NC * ND; k += 256)
    s_slot[k] = ((size_t)c0 * ND + k < (size_t)n_cells * ND) ? g.slot[(size_t)c0 * ND + k] : -1;
  __syncthreads();
  // E = alpha R on the matrix cores
  const int wv = tid >> 6, lane = tid & 63, r16 = lane & 15, kq = lane >> 4;
  T a[4][8];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int s = 0; s < 8; ++s) a[mt][s] = (T)s_al[16 * mt + r16][4 * s + kq];
  for (int nt = wv; nt < kQ2N / 16; nt += 4) {
    const int li = nt >> 1, j = (nt & 1) * 16 + r16;
    const int my_slot = s_slot[lane * ND + li];
    T b[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) b[s] = R[(size_t)(4 * s + kq) * kQ2N + li * 32 + j];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      typename Acc16<T>::type acc = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < 8; ++s) acc = mfma16(a[mt][s], b[s], acc);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        // the padding entries (j >= 27, zero rows of R) are stored too, so
        // that every 128-byte line of a slot is written whole
        const int sl = __shfl(my_slot, 16 * mt + mfma16_row(T(), u, lane));
        if (sl >= 0) E[(size_t)sl * 32 + j] = acc[u];
      }
    }
  }
}

// Persistent row gather: every half-wave streams a contiguous chunk of
// active rows.  For each row the element rows of its slots (lanes 0-26),
// their positions and (when adding) the row's old values are loaded, the
// element rows are added into the row's LDS copy in slot (cell) order, and the
// row is written with coalesced stores.  The next row's pointers and data are
// loaded before the current row is accumulated, so one memory round trip per
// row is in flight while the previous one completes.  (FP32, overwrite, 64^3:
// 0.65 ms; one short-lived workgroup per 8 rows 0.79 ms, two rows per
// half-wave or an XCD-ordered row mapping slower still.)
constexpr int kGatherBlocks = 1280;   // 5 per CU: all resident at the kernel's occupancy

template <typename T>
struct GatherRow {
  int rs = 0, len = 0, k0 = 0, nk = 0;
  T v[kQ2MaxInc];
  int o[kQ2MaxInc];
  double old[kQ2MaxRow / 32];
  double fv = 0.0, rold = 0.0;   // load-vector entry of slot l; the row's old rhs (lane 0)
};

template <typename T, bool OVERWRITE>
__device__ __forceinline__ void gather_ptrs(const Q2Gather& g, const DevCSR& A, int row, int row_end, GatherRow<T>& w) {
  w.rs = w.len = w.k0 = w.nk = 0;
  if (row < row_end) {
    w.rs = A.row_ptr[row];
    w.len = A.row_ptr[row + 1] - w.rs;
    w.k0 = g.inc_ptr[row];
    w.nk = g.inc_ptr[row + 1] - w.k0;
  }
}

template <typename T, bool OVERWRITE>
__device__ __forceinline__ void gather_data(const Q2Gather& g, const DevCSR& A, const T* __restrict__ E,
                                            const double* __restrict__ fe, const double* __restrict__ rhs,
                                            int row, int l, GatherRow<T>& w) {
#pragma unroll
  for (int k = 0; k < kQ2MaxInc; ++k) {
    w.v[k] = 0;
    w.o[k] = 0;
    if (k < w.nk && l < 27) {
      w.v[k] = E[(size_t)(w.k0 + k) * 32 + l];
      w.o[k] = g.off[(size_t)(w.k0 + k) * 32 + l];
    }
  }
#pragma unroll
  for (int q = 0; q < kQ2MaxRow / 32; ++q)
    w.old[q] = (!OVERWRITE && 32 * q + l < w.len) ? A.val[w.rs + 32 * q + l] : 0.0;
  w.fv = (l < w.nk) ? fe[w.k0 + l] : 0.0;
  w.rold = (!OVERWRITE && l == 0 && w.len > 0) ? rhs[row] : 0.0;
}

template <typename T, bool OVERWRITE>
__global__ __launch_bounds__(256) void k_q2_row_gather(Q2Gather g, const T* __restrict__ E,
                                                       const double* __restrict__ fe, DevCSR A,
                                                       double* __restrict__ rhs) {
  constexpr int NQ = kQ2MaxRow / 32;
  __shared__ double s_row[8][kQ2MaxRow];
  const int h = threadIdx.x >> 5, l = threadIdx.x & 31;
  const int n_hw = gridDim.x * 8, hw = blockIdx.x * 8 + h;
  const int chunk = (g.n_active + n_hw - 1) / n_hw;
  const int r0 = min(hw * chunk, g.n_active), r1 = min(r0 + chunk, g.n_active);
  double* b = s_row[h];
  GatherRow<T> cur, nxt;
  gather_ptrs<T, OVERWRITE>(g, A, r0, r1, cur);
  gather_data<T, OVERWRITE>(g, A, E, fe, rhs, r0, l, cur);
  gather_ptrs<T, OVERWRITE>(g, A, r0 + 1, r1, nxt);
  for (int row = r0; row < r1; ++row) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) b[32 * q + l] = 0.0;
    wave_sync();
    // LDS adds of one wave run in issue order: each entry is summed in slot
    // order without a wait between the slots (the 27 positions of a slot
    // are distinct)
#pragma unroll
    for (int k = 0; k < kQ2MaxInc; ++k)
      if (k < cur.nk && l < 27) atomicAdd(&b[cur.o[k]], (double)cur.v[k]);
    // the next row's data and the pointers of the one after it
    GatherRow<T> nn;
    gather_data<T, OVERWRITE>(g, A, E, fe, rhs, row + 1, l, nxt);
    gather_ptrs<T, OVERWRITE>(g, A, row + 2, r1, nn);
    wave_sync();
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (32 * q + l < cur.len) A.val[cur.rs + 32 * q + l] = OVERWRITE ? b[32 * q + l] : cur.old[q] + b[32 * q + l];
    // the row's load vector: its slots' entries in slot order
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kQ2MaxInc; ++k) {
      const double t = __shfl(cur.fv, k, 32);
      if (k < cur.nk) s += t;
    }
    if (l == 0) rhs[row] = OVERWRITE ? s : cur.rold + s;
    wave_sync();
    cur = nxt;
    nxt = nn;
  }
}

__global__ void k_q2_gather_offsets(Q2Gather g, const int* __restrict__ inc, const int* __restrict__ dofs, DevCSR A,
                                    uint8_t* __restrict__ off, int* __restrict__ bad) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)g.n_inc * 32) return;
  const int k = (int)(t >> 5), j = (int)(t & 31);
  int o = 0;
  if (j < 27) {
    const int ic = inc[k], e = ic / 27, i = ic - 27 * e;
    const int r = dofs[(size_t)e * 27 + i], c = dofs[(size_t)e * 27 + j];
    const int rs = A.row_ptr[r], re = A.row_ptr[r + 1];
    const int p = find_col(A.col, rs, re, c);
    if (p >= re || A.col[p] != c || p - rs >= kQ2MaxRow) *bad = 1;
    else o = p - rs;
  }
  off[t] = (uint8_t)o;
}

__global__ void k_dirichlet_unit_diagonal(DevCSR A) {
  const int i = A.active + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A.n) A.val[A.row_ptr[i]] = 1.0;
}

template <typename T>
void q2_element_gemm_t(const AssemblyData& ad, const CDCoeffs& co, const Q2Gather& g, const T* R, T* E,
                       double* fe, hipStream_t st) {
  if (ad.n_cells > 0)
    hipLaunchKernelGGL(k_q2_element_gemm<T>, dim3((ad.n_cells + kQ2GemmCells - 1) / kQ2GemmCells), dim3(256), 0, st,
                       ad, co, g, R, E, fe);
}

template <typename T>
void q2_row_gather_t(const Q2Gather& g, const T* E, const double* fe, const DevCSR& A, double* rhs, bool overwrite,
                     hipStream_t st) {
  if (g.n_active > 0) {
    const int grid = std::min(kGatherBlocks, (g.n_active + 7) / 8);
    if (overwrite)
      hipLaunchKernelGGL((k_q2_row_gather<T, true>), dim3(grid), dim3(256), 0, st, g, E, fe, A, rhs);
    else
      hipLaunchKernelGGL((k_q2_row_gather<T, false>), dim3(grid), dim3(256), 0, st, g, E, fe, A, rhs);
  }
  if (A.n > A.active)
    hipLaunchKernelGGL(k_dirichlet_unit_diagonal, dim3((A.n - A.active + 255) / 256), dim3(256), 0, st, A);
}

}  // namespace

void q2_gather_offsets(const Q2Gather& g, const int* inc, const int* dofs, const DevCSR& A, uint8_t* off, int* bad,
                       hipStream_t st) {
  const size_t total = (size_t)g.n_inc * 32;
  if (total > 0)
    hipLaunchKernelGGL(k_q2_gather_offsets, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, g, inc, dofs,
                       A, off, bad);
}

void q2_element_gemm(const AssemblyData& ad, const CDCoeffs& co, const Q2Gather& g, const double* R, double* E,
                     double* fe, hipStream_t st) {
  q2_element_gemm_t(ad, co, g, R, E, fe, st);
}
void q2_element_gemm(const AssemblyData& ad, const CDCoeffs& co, const Q2Gather& g, const float* R, float* E,
                     double* fe, hipStream_t st) {
  q2_element_gemm_t(ad, co, g, R, E, fe, st);
}
void q2_row_gather(const Q2Gather& g, const double* E, const double* fe, const DevCSR& A, double* rhs,
                   bool overwrite, hipStream_t st) {
  q2_row_gather_t(g, E, fe, A, rhs, overwrite, st);
}
void q2_row_gather(const Q2Gather& g, const float* E, const double* fe, const DevCSR& A, double* rhs,
                   bool overwrite, hipStream_t st) {
  q2_row_gather_t(g, E, fe, A, rhs, overwrite, st);
}

void assemble_cd(const AssemblyData& ad, const CDCoeffs& co, const int* cells, int n,
                 const DevCSR& A, double* rhs, hipStream_t st) {
  if (ad.dim == 3 && ad.ndof == 27) {
    if (ad.nq <= 64) {
      if (n > 0)
        // the coefficient table of 32 quadrature points (Gauss3: 27) leaves
        // LDS for more workgroups per CU than the one of 64
        if (ad.nq <= 32)
          hipLaunchKernelGGL(k_assemble_cd_q2_mfma<32>, dim3((n + 3) / 4), dim3(256), 0, st, ad, co, cells, n, A, rhs);
        else
          hipLaunchKernelGGL(k_assemble_cd_q2_mfma<64>, dim3((n + 3) / 4), dim3(256), 0, st, ad, co, cells, n, A, rhs);
    } else {
      launch_asm<27, 3, 32>(ad, co, cells, n, A, rhs, st);
    }
  } else if (ad.dim == 3 && ad.ndof == 8) launch_asm<8, 3, 8>(ad, co, cells, n, A, rhs, st);
  else if (ad.dim == 2 && ad.ndof == 4) launch_asm<4, 2, 4>(ad, co, cells, n, A, rhs, st);
  else if (ad.dim == 2 && ad.ndof == 9) launch_asm<9, 2, 16>(ad, co, cells, n, A, rhs, st);
  else if (ad.dim == 2 && ad.ndof == 3) launch_asm<3, 2, 4>(ad, co, cells, n, A, rhs, st);
}

void dirichlet_rhs_example(int example, int n_dir, const double* dir_xyz, double* rhs_dir,
                           const CDCoeffs& co, hipStream_t st) {
  if (n_dir > 0)
    hipLaunchKernelGGL(k_dirichlet_example, dim3((n_dir + 255) / 256), dim3(256), 0, st, example,
                       co.time, n_dir, dir_xyz, rhs_dir);
}

void interpolate_example(int example, double time, int n, const double* xyz, double* u, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_interpolate_example, dim3((n + 255) / 256), dim3(256), 0, st, example, time, n,
                       xyz, u);
}

void assemble_tcd(const AssemblyData& ad, const CDCoeffs& co, const int* cells, int n, int active,
                  const DevCSR& A, const DevCSR& M, double* rhs, hipStream_t st) {
  if (ad.dim == 3 && ad.ndof == 27) launch_tcd<27, 3, 32>(ad, co, cells, n, active, A, M, rhs, st);
  else if (ad.dim == 3 && ad.ndof == 8) launch_tcd<8, 3, 8>(ad, co, cells, n, active, A, M, rhs, st);
  else if (ad.dim == 2 && ad.ndof == 4) launch_tcd<4, 2, 4>(ad, co, cells, n, active, A, M, rhs, st);
  else if (ad.dim == 2 && ad.ndof == 9) launch_tcd<9, 2, 16>(ad, co, cells, n, active, A, M, rhs, st);
  else if (ad.dim == 2 && ad.ndof == 3) launch_tcd<3, 2, 4>(ad, co, cells, n, active, A, M, rhs, st);
}

void cd_errors(const AssemblyData& ad, const CDCoeffs& co, const double* u, double* partial,
               double* out4, hipStream_t st) {
  const int nb = (ad.n_cells + 255) / 256;
  if (ad.dim == 3 && ad.ndof == 27)
    hipLaunchKernelGGL((k_cd_errors<27, 3>), dim3(nb), dim3(256), 0, st, ad, co, u, partial);
  else if (ad.dim == 3 && ad.ndof == 8)
    hipLaunchKernelGGL((k_cd_errors<8, 3>), dim3(nb), dim3(256), 0, st, ad, co, u, partial);
  else if (ad.dim == 3 && ad.ndof == 4)   // discontinuous P1 on hexahedra
    hipLaunchKernelGGL((k_cd_errors<4, 3>), dim3(nb), dim3(256), 0, st, ad, co, u, partial);
  else if (ad.dim == 2 && ad.ndof == 4)
    hipLaunchKernelGGL((k_cd_errors<4, 2>), dim3(nb), dim3(256), 0, st, ad, co, u, partial);
  else if (ad.dim == 2 && ad.ndof == 9)
    hipLaunchKernelGGL((k_cd_errors<9, 2>), dim3(nb), dim3(256), 0, st, ad, co, u, partial);
  else if (ad.dim == 2 && ad.ndof == 3)
    hipLaunchKernelGGL((k_cd_errors<3, 2>), dim3(nb), dim3(256), 0, st, ad, co, u, partial);
  else
    throw std::runtime_error("error norms: no kernel for this element");
  hipLaunchKernelGGL(k_reduce_errors, dim3(1), dim3(1024), 0, st, nb, partial, out4);
}

// ------------------------------------------ error norms without allocation
// Fixed device scratch of the error-norm reduction for callers that cannot
// allocate: the per-block partial sums (4 doubles per block of 256 cells, at
// most kErrorsScratchBlocks blocks) followed by the 4 results.  Every GPU the
// library runs on holds its own copy until the process ends; its address on a
// device comes from hipGetSymbolAddress with that device current (the name
// itself is only a host-side handle and must never reach a kernel).
static __device__ __align__(256) double scratch[(size_t)kErrorsScratchBlocks * 4 + 8];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

int errors_scratch_init(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) return (int)(e != hipSuccess ? e : hipErrorInvalidSymbol);
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return 0;
}

double* errors_scratch(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_ptr[device];
}

bool cd_errors_scratch(const AssemblyData& ad, const CDCoeffs& co, const double* u, double* scratch_ptr,
                       const double** out4, hipStream_t st) {
  const int nb = (ad.n_cells + 255) / 256;
  if (nb > kErrorsScratchBlocks || scratch_ptr == nullptr) return false;
  double* res = scratch_ptr + (size_t)kErrorsScratchBlocks * 4;
  cd_errors(ad, co, u, scratch_ptr, res, st);
  *out4 = res;
  return true;
}

}  // namespace pmx
