// Kernels of the batched time-dependent solve: many independent TCD problems
// (own mesh, coefficients and size each) advanced by backward Euler together.
//
// Per problem and time step this computes what TimeConvectionDiffusion<2>
// does per instance (src/System/TimeConvectionDiffusion.C:294-386 with
// TimeDiscretization::prepare_rhs_from_time_disc / prepare_system_matrix,
// src/TimeDiscRoutines/TimeDiscretizations.C:106-210):
//   rhs_td = tau f(t) + M u_old on active rows, rhs_td = g(t) on Dirichlet rows
//   (tau A + M) u = rhs_td,   u starting from u_old with the Dirichlet values.
// The linear solve is an iterative refinement: the FP64 residual of the FP64
// system matrix drives restarted GMRES(m) corrections computed in FP32 (FP32
// matrix and Krylov vectors, FP64 dot accumulation), right-preconditioned by
// Jacobi, until the FP64 residual is reduced by `red` (or below `tol`).
//
// Layout: the problems' DOFs are concatenated (problem p owns the global
// range [dof_off[p], dof_off[p+1]) in its own numbering: active DOFs first,
// Dirichlet DOFs last).  The active rows of every problem are cut into its own
// SELL slices of 64 rows (local column indices), so one workgroup can run the
// whole solve of one problem with workgroup barriers only: the solve kernel is
// one persistent launch with one workgroup per problem (largest first).
#include "common.hpp"
#include "examples.hpp"
#include "launch.hpp"

namespace pmx {

namespace {

constexpr int kMaxRestart = 64;

// sum over the workgroup, returned to every thread (fixed summation order)
template <int BS>
__device__ __forceinline__ double wg_sum(double v, double* red) {
  v = group_sum<kWave>(v);
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0;
#pragma unroll
  for (int i = 0; i < BS / kWave; ++i) s += red[i];
  __syncthreads();
  return s;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_batch_solve(BatchDev B, BatchSolveParams P) {
  constexpr int NW = BS / kWave;
  constexpr int LD = kMaxRestart + 1;
  __shared__ double red[NW];
  __shared__ double H[LD * kMaxRestart];
  __shared__ double cs[kMaxRestart], sn[kMaxRestart], g[LD], y[kMaxRestart];
  __shared__ double bc[1];
  const int p = B.order[blockIdx.x];
  const int off = B.dof_off[p], na = B.n_active[p];
  const int s0 = B.slice_off[p], s1 = B.slice_off[p + 1];
  double* x = B.x + off;
  const double* rhs = B.rhs + off;
  double* r = B.r + off;
  float* w = B.w + off;
  const float* dinv = B.dinv32 + off;
  float* V = B.V + B.ws_off[p];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int m = P.m;
  double rn0 = 0, rn = 0;
  int outer = 0, inner_total = 0;
  for (;;) {
    // FP64 residual r = rhs - S x on the active rows
    double part = 0;
    for (int s = s0 + wv; s < s1; s += NW) {
      const int lr = B.srow[(size_t)s * kWave + lane];
      const int a = B.sptr[s], width = (B.sptr[s + 1] - a) / kWave;
      double acc = 0;
      for (int j = 0; j < width; ++j) {
        const int k = a + j * kWave + lane;
        acc += B.S64[k] * x[B.scol[k]];
      }
      if (lr >= 0) {
        const double d = rhs[lr] - acc;
        r[lr] = d;
        part += d * d;
      }
    }
    rn = sqrt(wg_sum<BS>(part, red));
    if (outer == 0) rn0 = rn;
    if (rn <= P.red * rn0 || rn <= P.tol || outer >= P.max_outer) break;
    ++outer;
    // inner GMRES(m) for S d = r, d0 = 0, right Jacobi preconditioning
    const double inv = 1.0 / rn;
    for (int i = threadIdx.x; i < na; i += BS) V[i] = (float)(r[i] * inv);
    if (threadIdx.x == 0) g[0] = rn;
    __syncthreads();
    int k = 0;
    for (int j = 0; j < m; ++j) {
      const float* vj = V + (size_t)j * na;
      float* vn = V + (size_t)(j + 1) * na;
      for (int s = s0 + wv; s < s1; s += NW) {
        const int lr = B.srow[(size_t)s * kWave + lane];
        const int a = B.sptr[s], width = (B.sptr[s + 1] - a) / kWave;
        float acc = 0;
        for (int jj = 0; jj < width; ++jj) {
          const int kk = a + jj * kWave + lane;
          const int c = B.scol[kk];
          const float z = (c < na) ? dinv[c] * vj[c] : 0.0f;   // correction is 0 on Dirichlet DOFs
          acc += B.S32[kk] * z;
        }
        if (lr >= 0) w[lr] = acc;
      }
      __syncthreads();
      // modified Gram-Schmidt
      for (int i = 0; i <= j; ++i) {
        const float* vi = V + (size_t)i * na;
        double d = 0;
        for (int t = threadIdx.x; t < na; t += BS) d += (double)w[t] * (double)vi[t];
        const double h = wg_sum<BS>(d, red);
        const float hf = (float)h;
        for (int t = threadIdx.x; t < na; t += BS) w[t] -= hf * vi[t];
        if (threadIdx.x == 0) H[i + j * LD] = h;
      }
      double d = 0;
      for (int t = threadIdx.x; t < na; t += BS) d += (double)w[t] * (double)w[t];
      const double hn = sqrt(wg_sum<BS>(d, red));
      const float ih = (hn > 0) ? (float)(1.0 / hn) : 0.0f;
      for (int t = threadIdx.x; t < na; t += BS) vn[t] = w[t] * ih;
      if (threadIdx.x == 0) {
        H[j + 1 + j * LD] = hn;
        for (int i = 0; i < j; ++i) {
          const double t0 = H[i + j * LD], t1 = H[i + 1 + j * LD];
          H[i + j * LD] = cs[i] * t0 + sn[i] * t1;
          H[i + 1 + j * LD] = -sn[i] * t0 + cs[i] * t1;
        }
        const double a0 = H[j + j * LD], a1 = H[j + 1 + j * LD];
        double c = 1, sv = 0;
        if (a1 != 0) {
          const double rr = sqrt(a0 * a0 + a1 * a1);
          c = a0 / rr;
          sv = a1 / rr;
        }
        cs[j] = c;
        sn[j] = sv;
        H[j + j * LD] = c * a0 + sv * a1;
        H[j + 1 + j * LD] = 0;
        g[j + 1] = -sv * g[j];
        g[j] = c * g[j];
        bc[0] = (hn > 0) ? fabs(g[j + 1]) : 0.0;
      }
      __syncthreads();
      k = j + 1;
      if (bc[0] <= P.inner_red * rn) break;
    }
    inner_total += k;
    if (threadIdx.x == 0) {
      for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int l = i + 1; l < k; ++l) s -= H[i + l * LD] * y[l];
        const double dd = H[i + i * LD];
        y[i] = (dd != 0) ? s / dd : 0.0;
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < na; t += BS) {
      double acc = 0;
      for (int i = 0; i < k; ++i) acc += y[i] * (double)V[(size_t)i * na + t];
      x[t] += (double)dinv[t] * acc;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    B.iters[p] += inner_total;
    B.outer[p] += outer;
    B.res[p] = (rn0 > 0) ? rn / rn0 : 0.0;
    if (!(rn <= P.red * rn0 || rn <= P.tol)) B.fail[p] += 1;
  }
}

// rhs_td = tau f + M u_old on the active rows of every problem (SELL slices)
__global__ __launch_bounds__(256) void k_batch_rhs(int n_slices, const int* __restrict__ slice_prob,
                                                   const int* __restrict__ dof_off,
                                                   const int* __restrict__ sptr,
                                                   const int* __restrict__ srow,
                                                   const int* __restrict__ scol,
                                                   const double* __restrict__ M64, double tau,
                                                   const double* __restrict__ f,
                                                   const double* __restrict__ u,
                                                   double* __restrict__ rhs) {
  const int lane = threadIdx.x & (kWave - 1);
  const int s = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) / kWave);
  if (s >= n_slices) return;
  const int lr = srow[(size_t)s * kWave + lane];
  if (lr < 0) return;
  const int off = dof_off[slice_prob[s]];
  const int a = sptr[s], width = (sptr[s + 1] - a) / kWave;
  double acc = 0;
  for (int j = 0; j < width; ++j) {
    const int k = a + j * kWave + lane;
    acc += M64[k] * u[off + scol[k]];
  }
  double v = tau * f[off + lr];
  v += acc;
  rhs[off + lr] = v;
}

// Dirichlet rows: rhs_td = f (boundary value), u = f (copy_nonactive)
__global__ void k_batch_dirichlet_copy(int n, const int* __restrict__ idx, const double* __restrict__ f,
                                       double* __restrict__ rhs, double* __restrict__ u) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    const int i = idx[k];
    rhs[i] = f[i];
    u[i] = f[i];
  }
}

// out[idx[k]] = boundary value at time *t
__global__ void k_batch_dirichlet_values(int ex, const double* __restrict__ t, int n,
                                         const int* __restrict__ idx, const double* __restrict__ xyz,
                                         double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    const int i = idx[k];
    out[i] = example_dirichlet(ex, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], *t);
  }
}

// S = tau A + M on active rows, S = A on Dirichlet rows
__global__ void k_batch_system(int n_rows, const int* __restrict__ rp, const uint8_t* __restrict__ row_active,
                               double tau, const double* __restrict__ A, const double* __restrict__ M,
                               double* __restrict__ S) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const bool act = row_active[r] != 0;
  for (int k = rp[r]; k < rp[r + 1]; ++k) {
    if (act) {
      double v = tau * A[k];
      v += M[k];
      S[k] = v;
    } else {
      S[k] = A[k];
    }
  }
}

__global__ void k_add_scalar(double* __restrict__ t, double dt) {
  if (blockIdx.x == 0 && threadIdx.x == 0) t[0] += dt;
}

__global__ void k_sell_gather(size_t total, const int* __restrict__ src, const double* __restrict__ v,
                              double* __restrict__ out64, float* __restrict__ out32) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (size_t)gridDim.x * blockDim.x) {
    const int s = src[k];
    const double a = (s >= 0) ? v[s] : 0.0;
    if (out64) out64[k] = a;
    if (out32) out32[k] = (float)a;
  }
}

}  // namespace

void batch_solve(const BatchDev& B, const BatchSolveParams& P, int block, hipStream_t st) {
  switch (block) {
    case 256: hipLaunchKernelGGL((k_batch_solve<256>), dim3(B.n_prob), dim3(256), 0, st, B, P); break;
    case 512: hipLaunchKernelGGL((k_batch_solve<512>), dim3(B.n_prob), dim3(512), 0, st, B, P); break;
    default: hipLaunchKernelGGL((k_batch_solve<1024>), dim3(B.n_prob), dim3(1024), 0, st, B, P); break;
  }
}

void batch_rhs(int n_slices, const int* slice_prob, const int* dof_off, const int* sptr, const int* srow,
               const int* scol, const double* M64, double tau, const double* f, const double* u,
               double* rhs, hipStream_t st) {
  if (n_slices > 0)
    hipLaunchKernelGGL(k_batch_rhs, dim3((n_slices + 3) / 4), dim3(256), 0, st, n_slices, slice_prob, dof_off,
                       sptr, srow, scol, M64, tau, f, u, rhs);
}

void batch_dirichlet_copy(int n, const int* idx, const double* f, double* rhs, double* u, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_batch_dirichlet_copy, dim3((n + 255) / 256), dim3(256), 0, st, n, idx, f, rhs, u);
}

void batch_dirichlet_values(int example, const double* t_dev, int n, const int* idx, const double* xyz,
                            double* out, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_batch_dirichlet_values, dim3((n + 255) / 256), dim3(256), 0, st, example, t_dev, n,
                       idx, xyz, out);
}

void batch_system(int n_rows, const int* rp, const uint8_t* row_active, double tau, const double* A,
                  const double* M, double* S, hipStream_t st) {
  if (n_rows > 0)
    hipLaunchKernelGGL(k_batch_system, dim3((n_rows + 255) / 256), dim3(256), 0, st, n_rows, rp, row_active,
                       tau, A, M, S);
}

void add_scalar(double* t, double dt, hipStream_t st) {
  hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(64), 0, st, t, dt);
}

void sell_gather(size_t total, const int* src, const double* v, double* out64, float* out32, hipStream_t st) {
  if (total)
    hipLaunchKernelGGL(k_sell_gather, dim3(4096), dim3(256), 0, st, total, src, v, out64, out32);
}

}  // namespace pmx
