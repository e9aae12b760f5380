// Oseen / Navier-Stokes saddle-point kernels (gfx950): Q2/Q1 assembly of the
// NSTYPE-1 blocks, the combined matrix, the pressure mean, and the batched
// cell-Vanka smoother.
//
// Assembly restates, per hexahedron and quadrature point, the routines that
// LocalAssembling<3>::set_parameters_for_nse selects for a Galerkin NSTYPE 1
// problem with the convective form (src/AssembleRoutines/LocalAssembling.C:
// 931-1185, src/AssembleRoutinesSaddle/NSE_local_assembling_routines.C):
//   NSLaplaceGradGradSingle            A_ij += w|J| nu grad phi_j . grad phi_i
//   NSNonlinearTerm_convective_Single  A_ij += w|J| (w.grad phi_j) phi_i,
//                                      w = the given velocity field at the point
//   NSDivergenceBlocks (sign 1)        Bk_ij -= w|J| psi_i d_k phi_j
//   NSRightHandSide (sign -1)          f_k,i += w|J| f_k phi_i;  g_i += w|J| psi_i g
// with the scatter of Assemble3D (Dirichlet velocity rows of A get diagonal 1;
// B holds every velocity column).  Lanes: 32 per cell, lane i < 27 owns
// velocity test function i (its A row, its three load entries and the column
// i of B1..B3 for the 8 pressure rows); cells of one colour share no vertex, so
// neither velocity nor pressure rows are written twice.
#include <stdexcept>

#include "common.hpp"
#include "examples.hpp"
#include "launch.hpp"

namespace pmx {

namespace {

__device__ __forceinline__ int find_col_n(const int* __restrict__ ci, int s, int e, int c) {
  int lo = s, hi = e - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ci[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

constexpr int NU = 27, NP = 8, SLOT = 32, CPB = 256 / SLOT;
// reference table per quadrature point: w, xi[3], phi[27], dphi[3][27], psi[8]
constexpr int RS = 1 + 3 + NU + 3 * NU + NP;

// NPC: pressure DOFs per cell (8: Q1, 4: discontinuous P1); the table keeps 8
// psi slots either way
template <int NPC>
__global__ __launch_bounds__(256) void k_assemble_nse(NseAsm d, const int* __restrict__ cells, int n) {
  __shared__ double s_u[CPB][NU];
  __shared__ double s_g[CPB][3][NU];
  __shared__ double s_w[CPB][3][NU];
  __shared__ double s_c[CPB][8];
  const int slot = threadIdx.x / SLOT;
  const int i = threadIdx.x % SLOT;
  const int ci = blockIdx.x * CPB + slot;
  const bool valid = ci < n;
  const int cell = valid ? cells[ci] : 0;
  const double* geo = d.geo + (size_t)cell * kGeoStride;
  const int* vd = d.vdofs + (size_t)cell * NU;
  const int* pd = d.pdofs + (size_t)cell * NPC;
  double Jinv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Jinv[k] = geo[12 + k];
  if (valid && i < NU) {
    for (int k = 0; k < 3; ++k) s_w[slot][k][i] = d.w ? d.w[(size_t)k * d.n_u + vd[i]] : 0.0;
  }
  double arow[NU];
#pragma unroll
  for (int j = 0; j < NU; ++j) arow[j] = 0.0;
  double bcol[3][NPC];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int q = 0; q < NPC; ++q) bcol[k][q] = 0.0;
  double frow[3] = {0, 0, 0}, grow = 0;
  __syncthreads();
  for (int q = 0; q < d.nq; ++q) {
    const double* rq = d.ref + (size_t)q * RS;
    if (valid && i < NU) {
      const double* dphi = rq + 4 + NU;
      s_u[slot][i] = rq[4 + i];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) s += dphi[k * NU + i] * Jinv[k * 3 + a];
        s_g[slot][a][i] = s;
      }
    }
    __syncthreads();
    if (valid && i == 0) {
      double x[3];
      for (int r = 0; r < 3; ++r) {
        double s = geo[r];
        for (int c = 0; c < 3; ++c) s += geo[3 + r * 3 + c] * rq[1 + c];
        x[r] = s;
      }
      double f[4];
      nse_example_rhs(d.example, x[0], x[1], x[2], d.nu, f);
      double* sc = s_c[slot];
      sc[0] = f[0]; sc[1] = f[1]; sc[2] = f[2]; sc[3] = f[3];
      sc[4] = rq[0] * fabs(geo[21]);
      for (int k = 0; k < 3; ++k) {
        double s = 0;
        for (int j = 0; j < NU; ++j) s += s_w[slot][k][j] * s_u[slot][j];
        sc[5 + k] = s;
      }
    }
    __syncthreads();
    if (valid && i < NU) {
      const double* sc = s_c[slot];
      const double mult = sc[4], w1 = sc[5], w2 = sc[6], w3 = sc[7];
      const double ti = s_u[slot][i];
      const double tx = s_g[slot][0][i], ty = s_g[slot][1][i], tz = s_g[slot][2][i];
      const double nu = d.nu;
#pragma unroll
      for (int j = 0; j < NU; ++j) {
        const double ax = s_g[slot][0][j], ay = s_g[slot][1][j], az = s_g[slot][2][j];
        arow[j] += mult * nu * (tx * ax + ty * ay + tz * az);
        if (d.convection) arow[j] += mult * (w1 * ax + w2 * ay + w3 * az) * ti;
      }
      frow[0] += mult * ti * sc[0];
      frow[1] += mult * ti * sc[1];
      frow[2] += mult * ti * sc[2];
      const double* psi = rq + 4 + NU + 3 * NU;
#pragma unroll
      for (int p = 0; p < NPC; ++p) {
        bcol[0][p] -= mult * psi[p] * tx;
        bcol[1][p] -= mult * psi[p] * ty;
        bcol[2][p] -= mult * psi[p] * tz;
      }
      if (i < NPC) grow += mult * psi[i] * sc[3];
    }
    __syncthreads();
  }
  if (!valid || i >= NU) return;
  const int r = vd[i];
  if (r < d.n_u_active) {
    const int rs = d.A.row_ptr[r], re = d.A.row_ptr[r + 1];
#pragma unroll
    for (int j = 0; j < NU; ++j) d.A.val[find_col_n(d.A.col, rs, re, vd[j])] += arow[j];
    if (d.rhs)
      for (int k = 0; k < 3; ++k) d.rhs[(size_t)k * d.n_u + r] += frow[k];
  } else {
    d.A.val[find_col_n(d.A.col, d.A.row_ptr[r], d.A.row_ptr[r + 1], r)] = 1.0;
  }
#pragma unroll
  for (int p = 0; p < NPC; ++p) {
    const int pr = pd[p];
    const int pos = find_col_n(d.Bcol, d.Brp[pr], d.Brp[pr + 1], r);
    d.B1[pos] += bcol[0][p];
    d.B2[pos] += bcol[1][p];
    d.B3[pos] += bcol[2][p];
  }
  if (d.rhs && i < NPC) d.rhs[(size_t)3 * d.n_u + pd[i]] += grow;
}

// velocity Dirichlet rows of the rhs: the example's boundary values
__global__ void k_nse_dirichlet(int ex, int n_u, int active, const double* __restrict__ xyz,
                                double* __restrict__ rhs) {
  const int k = active + blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_u) {
    for (int c = 0; c < 3; ++c) {
      double v[6];
      example_exact(ex + c, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], v);
      rhs[(size_t)c * n_u + k] = v[0];
    }
  }
}

// combined matrix values: kind 0 -> 0, 1 -> 1, 2 -> A[pos], 3..5 -> B1..B3[pos]
__global__ void k_combined_values(long long nnz, const uint8_t* __restrict__ kind, const int* __restrict__ pos,
                                  const double* __restrict__ A, const double* __restrict__ B1,
                                  const double* __restrict__ B2, const double* __restrict__ B3,
                                  double* __restrict__ out) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (long long)gridDim.x * blockDim.x) {
    const int t = kind[k], p = pos[k];
    double v = 0.0;
    if (t == 1) v = 1.0;
    else if (t == 2) v = A[p];
    else if (t == 3) v = B1[p];
    else if (t == 4) v = B2[p];
    else if (t == 5) v = B3[p];
    out[k] = v;
  }
}

// integral and measure of the pressure over the cells (per-block partials);
// NPC pressure DOFs per cell
template <int NPC>
__global__ __launch_bounds__(256) void k_pressure_integral(int n_cells, int nq, const double* __restrict__ ref,
                                                           const double* __restrict__ geo,
                                                           const int* __restrict__ pdofs,
                                                           const double* __restrict__ p,
                                                           double* __restrict__ partial) {
  __shared__ double red[16];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double I = 0, m = 0;
  if (c < n_cells) {
    const double det = fabs(geo[(size_t)c * kGeoStride + 21]);
    double pl[NPC];
    for (int j = 0; j < NPC; ++j) pl[j] = p[pdofs[(size_t)c * NPC + j]];
    for (int q = 0; q < nq; ++q) {
      const double* rq = ref + (size_t)q * RS;
      const double* psi = rq + 4 + NU + 3 * NU;
      double v = 0;
      for (int j = 0; j < NPC; ++j) v += pl[j] * psi[j];
      I += rq[0] * det * v;
      m += rq[0] * det;
    }
  }
  const double si = block_sum(I, red);
  const double sm = block_sum(m, red);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = si;
    partial[2 * blockIdx.x + 1] = sm;
  }
}

__global__ __launch_bounds__(1024) void k_pressure_shift(int nb, const double* __restrict__ partial, int n_p,
                                                         double* __restrict__ p) {
  __shared__ double red[16];
  __shared__ double mean;
  double si = 0, sm = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) { si += partial[2 * b]; sm += partial[2 * b + 1]; }
  const double ti = block_sum(si, red);
  const double tm = block_sum(sm, red);
  if (threadIdx.x == 0) mean = (tm > 0) ? ti / tm : 0.0;
  __syncthreads();
  for (int k = threadIdx.x; k < n_p; k += blockDim.x) p[k] -= mean;
}

// ------------------------------------------------------------- cell Vanka
// VankaSmoother with CELL_JACOBI (src/Multigrid/VankaSmoother.C:44-300): one
// local system per cell on the DOFs [u1 | u2 | u3 | p] of the cell (sorted per
// block), taken from the combined matrix; LU with partial pivoting (the
// LAPACK dgetrf of DenseMatrix::decomposeLU, src/Matrix/DenseMatrix.C:136-182);
// one smoothing step: r = rhs - K x, every cell solves its local system with
// the local r, the damped local solutions are summed per DOF (in cell order)
// and divided by the number of cells that touched the DOF.
// A pressure element with 4 DOFs per cell (discontinuous P1) leaves local
// rows 85..88 without a global DOF: their map entry is kVankaNone (sorted
// last), the factorisation gives them a unit row and column, so they neither
// pivot nor couple and the 85 real rows are factored exactly as an 85 x 85
// system; their defects read as 0 and their updates are never accumulated.
constexpr int NL = 3 * NU + NP;   // 89
constexpr int LDL = NL + 1;       // LDS row stride

__device__ __forceinline__ double vk_defect(const double* __restrict__ r, int g) {
  return g != kVankaNone ? r[g] : 0.0;
}

// one workgroup (256 threads) per cell: gather, factor in LDS (FP64), store the
// factors column-major (FP64 and/or FP32) and, in place of the pivots, the
// global DOF whose defect lands in local row i after all row interchanges
// (the interchanges of dgetrs applied in factorisation order compose to one
// permutation, so the solve gathers the permuted defect directly)
// With `inverse`, the factors are turned into the inverse of the
// row-interchanged local matrix in place, as LAPACK dgetri does: U^-1 column
// by column (upper triangle), then X = U^-1 L^-1 column by column from the
// right; the stored matrix then maps the permuted local defect to the local
// solution by one product.
__global__ __launch_bounds__(256) void k_vanka_factor(int n_cells, const int* __restrict__ map,
                                                      const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const double* __restrict__ cv, double* __restrict__ lu64,
                                                      float* __restrict__ lu32, int* __restrict__ pmap,
                                                      int* __restrict__ singular, int inverse, int ld) {
  __shared__ double M[NL * LDL];
  __shared__ double tcol[NL];
  __shared__ int dm[NL];
  __shared__ int pk[NL];
  __shared__ int src[NL];
  __shared__ int sp;
  const int cell = blockIdx.x;
  if (cell >= n_cells) return;   // uniform for the workgroup
  const int t = threadIdx.x;
  for (int a = t; a < NL; a += blockDim.x) dm[a] = map[(size_t)cell * NL + a];
  for (int k = t; k < NL * LDL; k += blockDim.x) M[k] = 0.0;
  __syncthreads();
  // gather: each wavefront takes local rows in turn and its 64 lanes stride
  // the row's CSR entries (64 loads in flight instead of one thread walking a
  // row of up to a few hundred entries); every entry finds its local column
  // by binary search in the sorted map
  {
    const int lane = t & 63, wv = t >> 6, nwv = blockDim.x >> 6;
    for (int a = wv; a < NL; a += nwv) {
      const int r = dm[a];
      if (r == kVankaNone) {
        if (lane == 0) M[a * LDL + a] = 1.0;
        continue;
      }
      const int qe = rp[r + 1];
      for (int q = rp[r] + lane; q < qe; q += 64) {
        const int c = ci[q];
        const double v = cv[q];
        int lo = 0, hi = NL - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (dm[mid] < c) lo = mid + 1; else hi = mid;
        }
        if (dm[lo] == c) M[a * LDL + lo] = v;
      }
    }
  }
  __syncthreads();
  for (int k = 0; k < NL; ++k) {
    if (t < 64) {
      // pivot: largest |M[i][k]|, i >= k, first index on ties
      double best = -1.0;
      int bi = k;
      for (int i = k + t; i < NL; i += 64) {
        const double v = fabs(M[i * LDL + k]);
        if (v > best) { best = v; bi = i; }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      if (t == 0) sp = bi;
    }
    __syncthreads();
    const int p = sp;
    if (p != k)
      for (int j = t; j < NL; j += blockDim.x) {
        const double tmp = M[k * LDL + j];
        M[k * LDL + j] = M[p * LDL + j];
        M[p * LDL + j] = tmp;
      }
    __syncthreads();
    const double d = M[k * LDL + k];
    if (t == 0) {
      pk[k] = p;
      if (d == 0.0) atomicOr(singular, 1);
    }
    const double inv = (d != 0.0) ? 1.0 / d : 0.0;
    for (int i = k + 1 + t; i < NL; i += blockDim.x) M[i * LDL + k] *= inv;
    __syncthreads();
    const int m = NL - k - 1;
    for (int e = t; e < m * m; e += blockDim.x) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      M[i * LDL + j] -= M[i * LDL + k] * M[k * LDL + j];
    }
    __syncthreads();
  }
  if (inverse) {
    // U^-1 in the upper triangle: column j from the columns before it
    for (int j = 0; j < NL; ++j) {
      for (int k = t; k <= j; k += blockDim.x) tcol[k] = M[k * LDL + j];   // column j of U
      __syncthreads();
      const double djj = tcol[j];
      const double ij = (djj != 0.0) ? 1.0 / djj : 0.0;
      for (int i = t; i < j; i += blockDim.x) {
        double acc = 0.0;
#pragma unroll 8
        for (int k = i; k < j; ++k) acc += M[i * LDL + k] * tcol[k];
        M[i * LDL + j] = -acc * ij;
      }
      if (t == 0) M[j * LDL + j] = ij;
      __syncthreads();
    }
    // X = U^-1 L^-1: column j = (U^-1)(:, j) - sum_{k > j} X(:, k) L(k, j)
    for (int j = NL - 2; j >= 0; --j) {
      for (int k = j + 1 + t; k < NL; k += blockDim.x) tcol[k] = M[k * LDL + j];   // column j of L
      __syncthreads();
      for (int i = t; i < NL; i += blockDim.x) {
        double acc = (i <= j) ? M[i * LDL + j] : 0.0;
#pragma unroll 8
        for (int k = j + 1; k < NL; ++k) acc -= M[i * LDL + k] * tcol[k];
        M[i * LDL + j] = acc;
      }
      __syncthreads();
    }
  }
  if (t == 0) {
    // compose the interchanges: local row i of the permuted system holds the
    // defect of local DOF src[i]
    for (int i = 0; i < NL; ++i) src[i] = i;
    for (int k = 0; k < NL; ++k) {
      const int p = pk[k];
      const int a = src[k];
      src[k] = src[p];
      src[p] = a;
    }
  }
  __syncthreads();
  for (int i = t; i < NL; i += blockDim.x) pmap[(size_t)cell * NL + i] = dm[src[i]];
  // column-major store: element (i, j) at j * ld + i (ld = NL, or NL + 1
  // with a zero in the padding row)
  for (int e = t; e < NL * ld; e += blockDim.x) {
    const int j = e / ld, i = e % ld;
    const double v = i < NL ? M[i * LDL + j] : 0.0;
    if (lu64) lu64[(size_t)cell * NL * ld + e] = v;
    if (lu32) lu32[(size_t)cell * NL * ld + e] = (float)v;
  }
}

// value of v in lane `lane` (uniform) for every lane of the wavefront
__device__ __forceinline__ double lane_value(double v, int lane) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// forward and backward substitution with the LU factors L of one cell
// (column-major, unit lower part) for the local rhs held in registers (rows
// lane and lane + 64 of the row-interchanged system); on return b0/b1 hold the
// local solution.  The solved component is broadcast by a lane read; the factor
// columns of the next U steps are loaded before those steps run (two register
// buffers), since their addresses do not depend on the solution.  Lanes
// without a second row (lane + 64 >= 89) load a valid duplicate and ignore it,
// so no load is predicated.
template <typename T>
__device__ __forceinline__ void vanka_substitute(const T* __restrict__ L, int lane, double& b0, double& b1) {
  constexpr int U = 8;
  const int i0 = lane, i1 = lane + 64;
  const bool h1 = i1 < NL;
  const int j1 = h1 ? i1 : i0;
  // column blocks of U steps; two register buffers so that the loads of the
  // next block are in flight while the current block is computed
  static_assert((NL - 1) % U == 0, "column blocks of the forward sweep");
  // forward-sweep block: columns k0 .. k0+U-1 (all < NL-1)
  auto load_l = [&](int k0, T(&l0)[U], T(&l1)[U]) {
    const T* col = L + (size_t)k0 * NL;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      l0[u] = col[u * NL + i0];
      l1[u] = col[u * NL + j1];
    }
  };
  // backward-sweep block: columns k0, k0-1, ..., k0-U+1; the last block
  // (k0 < U-1) reads column 0 in place of the negative ones and skips them
  auto load_u = [&](int k0, T(&l0)[U], T(&l1)[U], T(&dg)[U]) {
    if (k0 >= U - 1) {
      const T* col = L + (size_t)k0 * NL;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        l0[u] = col[-u * NL + i0];
        l1[u] = col[-u * NL + j1];
        dg[u] = col[-u * NL + (k0 - u)];
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 - u >= 0 ? k0 - u : 0;
        const T* col = L + (size_t)k * NL;
        l0[u] = col[i0];
        l1[u] = col[j1];
        dg[u] = col[k];
      }
    }
  };
  // L y = b (unit diagonal), columns 0 .. NL-2
  auto lower = [&](int k0, const T(&l0)[U], const T(&l1)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u;
      if (k < NL - 1) {
        const double bk = lane_value(k < 64 ? b0 : b1, k & 63);
        if (i0 > k) b0 -= (double)l0[u] * bk;
        if (h1 && i1 > k) b1 -= (double)l1[u] * bk;
      }
    }
  };
  // U x = y, columns NL-1 .. 0
  auto upper = [&](int k0, const T(&l0)[U], const T(&l1)[U], const T(&dg)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        const double bk = lane_value(k < 64 ? b0 : b1, k & 63);
        const double xk = bk / (double)dg[u];
        if (i0 == k) b0 = xk;
        if (i1 == k) b1 = xk;
        if (i0 < k) b0 -= (double)l0[u] * xk;
        if (h1 && i1 < k) b1 -= (double)l1[u] * xk;
      }
    }
  };
  {
    T a0[U], a1[U], c0[U], c1[U];
    load_l(0, a0, a1);
#pragma unroll 1
    for (int k0 = 0; k0 < NL - 1; k0 += 2 * U) {
      if (k0 + U < NL - 1) load_l(k0 + U, c0, c1);
      lower(k0, a0, a1);
      if (k0 + 2 * U < NL - 1) load_l(k0 + 2 * U, a0, a1);
      if (k0 + U < NL - 1) lower(k0 + U, c0, c1);
    }
  }
  {
    T a0[U], a1[U], ad[U], c0[U], c1[U], cd[U];
    load_u(NL - 1, a0, a1, ad);
#pragma unroll 1
    for (int k0 = NL - 1; k0 >= 0; k0 -= 2 * U) {
      if (k0 - U >= 0) load_u(k0 - U, c0, c1, cd);
      upper(k0, a0, a1, ad);
      if (k0 - 2 * U >= 0) load_u(k0 - 2 * U, a0, a1, ad);
      if (k0 - U >= 0) upper(k0 - U, c0, c1, cd);
    }
  }
}

// one wavefront per cell: gather the permuted local defect, substitute, store
// the damped local update (Jacobi variant: summed per DOF afterwards)
template <typename T>
__global__ __launch_bounds__(256) void k_vanka_solve(int n_cells, const T* __restrict__ lu,
                                                     const int* __restrict__ pmap,
                                                     const double* __restrict__ r, double damp,
                                                     double* __restrict__ upd) {
  const int lane = threadIdx.x & 63;
  const int cell = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (cell >= n_cells) return;   // uniform for the wavefront
  const int* pm = pmap + (size_t)cell * NL;
  const int i0 = lane, i1 = lane + 64;
  const bool h1 = i1 < NL;
  const int j1 = h1 ? i1 : i0;
  double b0 = vk_defect(r, pm[i0]);
  double b1 = vk_defect(r, pm[j1]);
  if (!h1) b1 = 0.0;
  vanka_substitute<T>(lu + (size_t)cell * NL * NL, lane, b0, b1);
  upd[(size_t)cell * NL + i0] = damp * b0;
  if (h1) upd[(size_t)cell * NL + i1] = damp * b1;
}

__device__ __forceinline__ double vk_entry(double v) { return v; }
__device__ __forceinline__ double vk_entry(float v) { return (double)v; }

// the same with the explicit local inverse: x = Inv b for the permuted local
// defect b (column-major Inv; the 89 columns are independent products, so the
// loads of the next columns are in flight while the current ones are summed)
template <typename T>
__global__ __launch_bounds__(256) void k_vanka_apply(int n_cells, const T* __restrict__ inv,
                                                     const int* __restrict__ pmap,
                                                     const double* __restrict__ r, double damp,
                                                     double* __restrict__ upd) {
  constexpr int U = 8;
  const int lane = threadIdx.x & 63;
  const int cell = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (cell >= n_cells) return;   // uniform for the wavefront
  const int* pm = pmap + (size_t)cell * NL;
  const int i0 = lane, i1 = lane + 64;
  const bool h1 = i1 < NL;
  const int j1 = h1 ? i1 : i0;
  double b0 = vk_defect(r, pm[i0]);
  double b1 = vk_defect(r, pm[j1]);
  if (!h1) b1 = 0.0;
  const T* X = inv + (size_t)cell * NL * NL;
  double x0 = 0.0, x1 = 0.0;
  int j = 0;
  for (; j + U <= NL; j += U) {
    T c0[U], c1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      c0[u] = __builtin_nontemporal_load(X + (size_t)(j + u) * NL + i0);
      c1[u] = __builtin_nontemporal_load(X + (size_t)(j + u) * NL + j1);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = j + u;
      const double bk = lane_value(k < 64 ? b0 : b1, k & 63);
      x0 += vk_entry(c0[u]) * bk;
      x1 += vk_entry(c1[u]) * bk;
    }
  }
  for (; j < NL; ++j) {
    const double bk = lane_value(j < 64 ? b0 : b1, j & 63);
    x0 += vk_entry(X[(size_t)j * NL + i0]) * bk;
    x1 += vk_entry(X[(size_t)j * NL + j1]) * bk;
  }
  upd[(size_t)cell * NL + i0] = damp * x0;
  if (h1) upd[(size_t)cell * NL + i1] = damp * x1;
}

// bf16 local inverses with columns padded to NLP = 90 entries (4-byte
// aligned): lane l < 45 owns local rows 2l and 2l+1 and reads both entries of
// a column with one 4-byte load (89 loads per lane instead of 178); the
// padding row 89 is zero and never written back
constexpr int NLP = NL + 1;
__global__ __launch_bounds__(256) void k_vanka_apply_bf16p(int n_cells, const uint32_t* __restrict__ inv,
                                                           const int* __restrict__ pmap,
                                                           const double* __restrict__ r, double damp,
                                                           double* __restrict__ upd) {
  constexpr int U = 8, H = NLP / 2;
  const int lane = threadIdx.x & 63;
  const int cell = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (cell >= n_cells) return;   // uniform for the wavefront
  const int* pm = pmap + (size_t)cell * NL;
  const int i1 = lane + 64;
  const bool h1 = i1 < NL;
  double b0 = vk_defect(r, pm[lane]);
  double b1 = vk_defect(r, pm[h1 ? i1 : lane]);
  if (!h1) b1 = 0.0;
  const int q = lane < H ? lane : H - 1;
  const uint32_t* X = inv + (size_t)cell * (NLP / 2) * NL;   // NL columns of H words
  double x0 = 0.0, x1 = 0.0;
  int j = 0;
  for (; j + U <= NL; j += U) {
    uint32_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(X + (size_t)(j + u) * H + q);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = j + u;
      const double bk = lane_value(k < 64 ? b0 : b1, k & 63);
      x0 += (double)__uint_as_float(c[u] << 16) * bk;
      x1 += (double)__uint_as_float(c[u] & 0xFFFF0000u) * bk;
    }
  }
  for (; j < NL; ++j) {
    const uint32_t c = X[(size_t)j * H + q];
    const double bk = lane_value(j < 64 ? b0 : b1, j & 63);
    x0 += (double)__uint_as_float(c << 16) * bk;
    x1 += (double)__uint_as_float(c & 0xFFFF0000u) * bk;
  }
  if (lane < H) {
    upd[(size_t)cell * NL + 2 * lane] = damp * x0;
    if (2 * lane + 1 < NL) upd[(size_t)cell * NL + 2 * lane + 1] = damp * x1;
  }
}

// the same for FP64 / FP32 inverses in the padded layout: one 16- / 8-byte
// load per lane and column
template <typename T>
__global__ __launch_bounds__(256) void k_vanka_apply_p(int n_cells, const T* __restrict__ inv,
                                                       const int* __restrict__ pmap, const double* __restrict__ r,
                                                       double damp, double* __restrict__ upd) {
  typedef T pair_t __attribute__((ext_vector_type(2)));
  constexpr int U = 8, H = NLP / 2;
  const int lane = threadIdx.x & 63;
  const int cell = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (cell >= n_cells) return;   // uniform for the wavefront
  const int* pm = pmap + (size_t)cell * NL;
  const int i1 = lane + 64;
  const bool h1 = i1 < NL;
  double b0 = vk_defect(r, pm[lane]);
  double b1 = vk_defect(r, pm[h1 ? i1 : lane]);
  if (!h1) b1 = 0.0;
  const int q = lane < H ? lane : H - 1;
  const pair_t* X = reinterpret_cast<const pair_t*>(inv + (size_t)cell * NLP * NL);
  double x0 = 0.0, x1 = 0.0;
  int j = 0;
  for (; j + U <= NL; j += U) {
    pair_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = __builtin_nontemporal_load(X + (size_t)(j + u) * H + q);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = j + u;
      const double bk = lane_value(k < 64 ? b0 : b1, k & 63);
      x0 += (double)c[u].x * bk;
      x1 += (double)c[u].y * bk;
    }
  }
  for (; j < NL; ++j) {
    const pair_t c = X[(size_t)j * H + q];
    const double bk = lane_value(j < 64 ? b0 : b1, j & 63);
    x0 += (double)c.x * bk;
    x1 += (double)c.y * bk;
  }
  if (lane < H) {
    upd[(size_t)cell * NL + 2 * lane] = damp * x0;
    if (2 * lane + 1 < NL) upd[(size_t)cell * NL + 2 * lane + 1] = damp * x1;
  }
}

// FP32 local inverses (cell-major, columns of NL) to the padded bf16 layout
__global__ void k_inv_to_bf16p(long long n_out, const float* __restrict__ a, uint16_t* __restrict__ b) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n_out;
       t += (long long)gridDim.x * blockDim.x) {
    const long long col = t / NLP;   // cell * NL + j
    const int i = (int)(t - col * NLP);
    uint16_t h = 0;
    if (i < NL) {
      uint32_t u = __float_as_uint(a[col * NL + i]);
      u += 0x7FFFu + ((u >> 16) & 1u);
      h = (uint16_t)(u >> 16);
    }
    b[t] = h;
  }
}

// FP32 values rounded to bf16 (nearest even)
__global__ void k_f32_to_bf16(long long n, const float* __restrict__ a, uint16_t* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    uint32_t u = __float_as_uint(a[i]);
    u += 0x7FFFu + ((u >> 16) & 1u);
    b[i] = (uint16_t)(u >> 16);
  }
}

// VankaSmoother CELL (multiplicative) over the cells of one colour: cells of a
// colour share no DOF, so each wave computes its local defect rows
// f_i - sum_k K_ik x_k with the current x (row entries in CSR order, as the
// reference's loop), substitutes and adds the damped local solution to x.
// Colours run one after another, each seeing the updates of the previous ones.
template <typename T>
__global__ __launch_bounds__(256) void k_vanka_colour(int n, const int* __restrict__ cells,
                                                      const T* __restrict__ lu, const int* __restrict__ pmap,
                                                      const int* __restrict__ map, const int* __restrict__ rp,
                                                      const int* __restrict__ ci, const double* __restrict__ cv,
                                                      const double* __restrict__ f, double damp,
                                                      double* __restrict__ x) {
  const int lane = threadIdx.x & 63;
  const int w = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
  if (w >= n) return;   // uniform for the wavefront
  const int cell = cells[w];
  const int* pm = pmap + (size_t)cell * NL;
  const int i0 = lane, i1 = lane + 64;
  const bool h1 = i1 < NL;
  const int j1 = h1 ? i1 : i0;
  auto defect = [&](int g) {
    if (g == kVankaNone) return 0.0;
    double t = f[g];
    const int e = rp[g + 1];
    for (int k = rp[g]; k < e; ++k) t -= cv[k] * x[ci[k]];
    return t;
  };
  double b0 = defect(pm[i0]);
  double b1 = defect(pm[j1]);
  if (!h1) b1 = 0.0;
  vanka_substitute<T>(lu + (size_t)cell * NL * NL, lane, b0, b1);
  const int* dm = map + (size_t)cell * NL;
  if (dm[i0] != kVankaNone) x[dm[i0]] += damp * b0;
  if (h1 && dm[i1] != kVankaNone) x[dm[i1]] += damp * b1;
}

// x[i] += (sum of the cell updates of DOF i, in cell order) / their number
__global__ void k_vanka_accumulate(int n, const int* __restrict__ acc_ptr, const int* __restrict__ acc_idx,
                                   const double* __restrict__ upd, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = acc_ptr[i], e = acc_ptr[i + 1];
  if (a == e) return;
  double s = 0.0;
  for (int k = a; k < e; ++k) s += upd[acc_idx[k]];
  x[i] += s / (double)(e - a);
}

}  // namespace

void assemble_nse(const NseAsm& d, const int* cells, int n, hipStream_t st) {
  const int grid = (n + CPB - 1) / CPB;
  if (d.np != 8 && d.np != 4) throw std::runtime_error("pressure element with 8 or 4 DOFs per cell expected");
  if (grid <= 0) return;
  if (d.np == 4) hipLaunchKernelGGL(k_assemble_nse<4>, dim3(grid), dim3(256), 0, st, d, cells, n);
  else hipLaunchKernelGGL(k_assemble_nse<8>, dim3(grid), dim3(256), 0, st, d, cells, n);
}

void nse_dirichlet_rhs(int example, int n_u, int active, const double* xyz, double* rhs, hipStream_t st) {
  const int nd = n_u - active;
  if (nd > 0)
    hipLaunchKernelGGL(k_nse_dirichlet, dim3((nd + 255) / 256), dim3(256), 0, st, example, n_u, active, xyz, rhs);
}

void combined_values(long long nnz, const uint8_t* kind, const int* pos, const double* A, const double* B1,
                     const double* B2, const double* B3, double* out, hipStream_t st) {
  if (nnz > 0)
    hipLaunchKernelGGL(k_combined_values, dim3(4096), dim3(256), 0, st, nnz, kind, pos, A, B1, B2, B3, out);
}

void vanka_factor(int n_cells, const int* map, const DevCSR& C, double* lu64, float* lu32, int* pmap,
                  int* singular, hipStream_t st, bool inverse, bool padded) {
  if (n_cells > 0)
    hipLaunchKernelGGL(k_vanka_factor, dim3(n_cells), dim3(256), 0, st, n_cells, map, C.row_ptr, C.col, C.val,
                       lu64, lu32, pmap, singular, inverse ? 1 : 0, padded ? NLP : NL);
}

void vanka_apply_inverse(int n_cells, const double* inv64, const float* inv32, const int* pmap, const double* r,
                         double damp, double* upd, hipStream_t st) {
  if (n_cells <= 0) return;
  const int grid = (n_cells + 3) / 4;
  if (inv32) hipLaunchKernelGGL((k_vanka_apply<float>), dim3(grid), dim3(256), 0, st, n_cells, inv32, pmap, r,
                                damp, upd);
  else hipLaunchKernelGGL((k_vanka_apply<double>), dim3(grid), dim3(256), 0, st, n_cells, inv64, pmap, r, damp,
                          upd);
}

void vanka_apply_inverse_bf16(int n_cells, const uint16_t* inv16, const int* pmap, const double* r, double damp,
                              double* upd, hipStream_t st) {
  if (n_cells <= 0) return;
  hipLaunchKernelGGL(k_vanka_apply_bf16p, dim3((n_cells + 3) / 4), dim3(256), 0, st, n_cells,
                     reinterpret_cast<const uint32_t*>(inv16), pmap, r, damp, upd);
}

size_t vanka_bf16_inverse_elems(int n_cells) { return (size_t)n_cells * NL * NLP; }
size_t vanka_padded_inverse_elems(int n_cells) { return (size_t)n_cells * NL * NLP; }

void vanka_apply_inverse_padded(int n_cells, const double* inv64, const float* inv32, const int* pmap,
                                const double* r, double damp, double* upd, hipStream_t st) {
  if (n_cells <= 0) return;
  const int grid = (n_cells + 3) / 4;
  if (inv32) hipLaunchKernelGGL(k_vanka_apply_p<float>, dim3(grid), dim3(256), 0, st, n_cells, inv32, pmap, r, damp,
                                upd);
  else hipLaunchKernelGGL(k_vanka_apply_p<double>, dim3(grid), dim3(256), 0, st, n_cells, inv64, pmap, r, damp,
                          upd);
}

void vanka_inverse_to_bf16(int n_cells, const float* inv32, uint16_t* inv16, hipStream_t st) {
  const long long n = (long long)vanka_bf16_inverse_elems(n_cells);
  if (n > 0) hipLaunchKernelGGL(k_inv_to_bf16p, dim3(4096), dim3(256), 0, st, n, inv32, inv16);
}

void f32_to_bf16(long long n, const float* a, uint16_t* b, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_f32_to_bf16, dim3(4096), dim3(256), 0, st, n, a, b);
}

void vanka_colour(int n, const int* cells, const double* lu64, const float* lu32, const int* pmap, const int* map,
                  const DevCSR& C, const double* f, double damp, double* x, hipStream_t st) {
  if (n <= 0) return;
  const int grid = (n + 3) / 4;
  if (lu32)
    hipLaunchKernelGGL((k_vanka_colour<float>), dim3(grid), dim3(256), 0, st, n, cells, lu32, pmap, map, C.row_ptr,
                       C.col, C.val, f, damp, x);
  else
    hipLaunchKernelGGL((k_vanka_colour<double>), dim3(grid), dim3(256), 0, st, n, cells, lu64, pmap, map, C.row_ptr,
                       C.col, C.val, f, damp, x);
}

void vanka_solve(int n_cells, const double* lu64, const float* lu32, const int* pmap, const double* r,
                 double damp, double* upd, hipStream_t st) {
  if (n_cells <= 0) return;
  const int grid = (n_cells + 3) / 4;
  if (lu32) hipLaunchKernelGGL((k_vanka_solve<float>), dim3(grid), dim3(256), 0, st, n_cells, lu32, pmap, r,
                               damp, upd);
  else hipLaunchKernelGGL((k_vanka_solve<double>), dim3(grid), dim3(256), 0, st, n_cells, lu64, pmap, r, damp,
                          upd);
}

void vanka_accumulate(int n, const int* acc_ptr, const int* acc_idx, const double* upd, double* x,
                      hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_vanka_accumulate, dim3((n + 255) / 256), dim3(256), 0, st, n, acc_ptr, acc_idx, upd, x);
}

void pressure_project_l20(int n_cells, int nq, int np_cell, const double* ref, const double* geo, const int* pdofs,
                          int n_p, double* p, double* partial, hipStream_t st) {
  const int nb = (n_cells + 255) / 256;
  if (np_cell == 4)
    hipLaunchKernelGGL(k_pressure_integral<4>, dim3(nb), dim3(256), 0, st, n_cells, nq, ref, geo, pdofs, p, partial);
  else if (np_cell == 8)
    hipLaunchKernelGGL(k_pressure_integral<8>, dim3(nb), dim3(256), 0, st, n_cells, nq, ref, geo, pdofs, p, partial);
  else
    throw std::runtime_error("pressure element with 8 or 4 DOFs per cell expected");
  hipLaunchKernelGGL(k_pressure_shift, dim3(1), dim3(1024), 0, st, nb, partial, n_p, p);
}

}  // namespace pmx
