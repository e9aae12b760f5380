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
// one persistent launch with one workgroup per problem (largest first), or
// with several for the costliest problems (each kernel below says how).
#include <stdexcept>

#include "common.hpp"
#include "examples.hpp"
#include "launch.hpp"

namespace pmx {

namespace {

constexpr int kMaxRestart = 64;
constexpr unsigned kBatchSpinLimit = 1u << 22;

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

// Slice loops of one wave over its share of a problem's slices, with U
// slices in flight (independent accumulators) to hide memory latency: a
// single workgroup has to stream the whole problem.
template <int U, typename VT, typename AT, typename Z, typename F>
__device__ __forceinline__ void wave_spmv(const BatchDev& B, const VT* __restrict__ val, int s0, int s1, int wv,
                                          int nw, int lane, Z z, F f) {
  for (int sb = s0 + wv; sb < s1; sb += U * nw) {
    int lr[U], a[U], wd[U];
    AT acc[U];
    int wmax = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = sb + u * nw;
      const bool ok = s < s1;
      lr[u] = ok ? B.srow[(size_t)s * kWave + lane] : -1;
      a[u] = ok ? B.sptr[s] : 0;
      wd[u] = ok ? (B.sptr[s + 1] - a[u]) / kWave : 0;
      acc[u] = 0;
      wmax = wd[u] > wmax ? wd[u] : wmax;
    }
    for (int j = 0; j < wmax; ++j) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j < wd[u]) {
          const int k = a[u] + j * kWave + lane;
          acc[u] += (AT)val[k] * z(B.scol[k]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (lr[u] >= 0) f(lr[u], acc[u]);
  }
}

template <int U, typename F>
__device__ __forceinline__ void wave_rows(const BatchDev& B, int s0, int s1, int wv, int nw, int lane, F f) {
  for (int sb = s0 + wv; sb < s1; sb += U * nw) {
    int lr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = sb + u * nw;
      lr[u] = (s < s1) ? B.srow[(size_t)s * kWave + lane] : -1;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (lr[u] >= 0) f(lr[u]);
  }
}

// Restarted GMRES(m) inside the refinement loop.  Orthogonalisation by
// classical Gram-Schmidt: the j+1 projections of the new vector are summed in
// one sweep over the basis (eight basis vectors per sweep of w) and reduced
// together, then subtracted in one pass that also sums |w|^2, so a step needs
// two reductions instead of two per basis vector.  Every lane owns the rows of
// its slices for all vector passes.  A problem may get G > 1 workgroups
// (B.wg_prob), as in the BiCGStab kernel below: each takes a contiguous share
// of the slices; the SpMVs start behind a barrier of the problem's
// workgroups, the sums add the workgroup partials in rank order (every
// workgroup gets the same Hessenberg column and takes the same branches), and
// thread 0 of every workgroup runs the same Givens and triangular solves.
template <int BS>
__global__ __launch_bounds__(BS) void k_batch_solve(BatchDev B, BatchSolveParams P) {
  constexpr int NW = BS / kWave;
  constexpr int LD = kMaxRestart + 1;
  constexpr int U = 4;
  __shared__ double red[NW];
  __shared__ double dred[kMaxRestart * NW];
  __shared__ double hs[kMaxRestart];
  __shared__ float hf[kMaxRestart];
  __shared__ double H[LD * kMaxRestart];
  __shared__ double cs[kMaxRestart], sn[kMaxRestart], g[LD], y[kMaxRestart];
  __shared__ double bc[1];
  int p, rk = 0, G = 1;
  if (B.wg_prob) {
    p = B.wg_prob[blockIdx.x];
    rk = B.wg_rank[blockIdx.x];
    G = B.groups[p];
  } else {
    p = B.order[blockIdx.x];
  }
  const int off = B.dof_off[p], na = B.n_active[p];
  const int ps0 = B.slice_off[p], ps1 = B.slice_off[p + 1];
  const int s0 = ps0 + (int)((long long)(ps1 - ps0) * rk / G);
  const int s1 = ps0 + (int)((long long)(ps1 - ps0) * (rk + 1) / G);
  unsigned epoch = 0;
  int nred = 0;
  auto gbar = [&]() {
    __syncthreads();
    if (G == 1) return;
    ++epoch;
    if (threadIdx.x == 0) {
      __threadfence();
      atomicAdd(B.bar + p, 1u);
      const unsigned target = epoch * (unsigned)G;
      unsigned spins = 0;
      while (__hip_atomic_load(B.bar + p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++spins > kBatchSpinLimit) {
          atomicOr(B.bar_timeout, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __threadfence();
    }
    __syncthreads();
  };
  // hs[0..n) = the problem's sums of the workgroup sums in hs[0..n)
  // (n <= kMaxRestart; two slot buffers alternate as in the BiCGStab kernel)
  auto sum_groups = [&](int n) {
    if (G == 1) return;
    const int buf = (nred++) & 1;
    double* slots = B.part + ((size_t)p * 2 + buf) * kBatchMaxGroups * kBatchPartStride;
    // thread 0 publishes all n values: the barrier's release is its own
    if (threadIdx.x == 0)
      for (int i = 0; i < n; ++i) slots[rk * kBatchPartStride + i] = hs[i];
    gbar();
    if ((int)threadIdx.x < n) {
      double t = 0;
      for (int q = 0; q < G; ++q)
        t += __hip_atomic_load(slots + q * kBatchPartStride + threadIdx.x, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      hs[threadIdx.x] = t;
    }
    __syncthreads();
  };
  // one sum over the problem
  auto sum1 = [&](double v) {
    v = group_sum<kWave>(v);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
#pragma unroll
      for (int i = 0; i < NW; ++i) t += red[i];
      hs[0] = t;
    }
    __syncthreads();
    sum_groups(1);
    const double r = hs[0];
    __syncthreads();
    return r;
  };
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
    // FP64 residual r = rhs - S x on the active rows (x of all rows)
    gbar();
    double part = 0;
    wave_spmv<U, double, double>(B, B.S64, s0, s1, wv, NW, lane, [&](int c) { return x[c]; },
                                 [&](int lr, double acc) {
                                   const double d = rhs[lr] - acc;
                                   r[lr] = d;
                                   part += d * d;
                                 });
    rn = sqrt(sum1(part));
    if (outer == 0) rn0 = rn;
    if (rn <= P.red * rn0 || rn <= P.tol || outer >= P.max_outer) break;
    ++outer;
    // inner GMRES(m) for S d = r, d0 = 0, right Jacobi preconditioning
    const double inv = 1.0 / rn;
    wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) { V[lr] = (float)(r[lr] * inv); });
    if (threadIdx.x == 0) g[0] = rn;
    int k = 0;
    for (int j = 0; j < m; ++j) {
      const float* vj = V + (size_t)j * na;
      float* vn = V + (size_t)(j + 1) * na;
      // w = S D^-1 v_j (v_j of all rows; the correction is 0 on Dirichlet DOFs)
      gbar();
      wave_spmv<U, float, float>(B, B.S32, s0, s1, wv, NW, lane,
                                 [&](int c) { return (c < na) ? dinv[c] * vj[c] : 0.0f; },
                                 [&](int lr, float acc) { w[lr] = acc; });
      // h_i = (w, v_i), i <= j (each lane reads back the rows it wrote)
      for (int i0 = 0; i0 <= j; i0 += 8) {
        double d[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) d[u] = 0.0;
        wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) {
          const double wt = (double)w[lr];
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (i0 + u <= j) d[u] += wt * (double)V[(size_t)(i0 + u) * na + lr];
        });
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double sd = group_sum<kWave>(d[u]);
          if (lane == 0 && i0 + u <= j) dred[(i0 + u) * NW + wv] = sd;
        }
      }
      __syncthreads();
      if ((int)threadIdx.x <= j) {
        double h = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) h += dred[threadIdx.x * NW + q];
        hs[threadIdx.x] = h;
      }
      __syncthreads();
      sum_groups(j + 1);
      if ((int)threadIdx.x <= j) {
        H[threadIdx.x + j * LD] = hs[threadIdx.x];
        hf[threadIdx.x] = (float)hs[threadIdx.x];
      }
      __syncthreads();
      // w -= sum_i h_i v_i, |w|^2
      double dn = 0;
      wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) {
        float wt = w[lr];
#pragma unroll 8
        for (int i = 0; i <= j; ++i) wt -= hf[i] * V[(size_t)i * na + lr];
        w[lr] = wt;
        dn += (double)wt * (double)wt;
      });
      const double hn = sqrt(sum1(dn));
      const float ih = (hn > 0) ? (float)(1.0 / hn) : 0.0f;
      wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) { vn[lr] = w[lr] * ih; });
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
        double sacc = g[i];
        for (int l = i + 1; l < k; ++l) sacc -= H[i + l * LD] * y[l];
        const double dd = H[i + i * LD];
        y[i] = (dd != 0) ? sacc / dd : 0.0;
      }
    }
    __syncthreads();
    wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) {
      double acc = 0;
#pragma unroll 8
      for (int i = 0; i < k; ++i) acc += y[i] * (double)V[(size_t)i * na + lr];
      x[lr] += (double)dinv[lr] * acc;
    });
  }
  if (threadIdx.x == 0 && rk == 0) {
    B.iters[p] += inner_total;
    B.outer[p] += outer;
    B.res[p] = (rn0 > 0) ? rn / rn0 : 0.0;
    if (!(rn <= P.red * rn0 || rn <= P.tol)) B.fail[p] += 1;
  }
}

// Same refinement loop with BiCGStab (Iteration_bicgstab's method,
// src/Solver/Iteration_bicgstab.C) as the FP32 inner solver, right-
// preconditioned by Jacobi.  Every lane owns the rows of its slices for all
// vector updates, so the only cross-thread data flow is the gather of the
// SpMVs; dots are fused into the passes that produce their operands.
// Workspace per problem: ri, rh, p, v, s, t, d (7 x n_active floats).
//
// A problem may get G > 1 workgroups (B.wg_prob): each takes a contiguous
// share of the problem's slices; the vectors stay the problem's, so the
// passes that read other rows (the SpMVs, the FP64 residual) are preceded by
// a barrier of the problem's workgroups, and every dot product adds the G
// workgroup partials in rank order, so all G see the same value and take
// the same branches.  The host launches such problems only when every
// workgroup of the grid can be resident at once; each barrier wait is
// bounded and sets B.bar_timeout when it gives up, so the kernel always ends.

template <int BS>
__global__ __launch_bounds__(BS) void k_batch_solve_bicgstab(BatchDev B, BatchSolveParams P) {
  constexpr int NW = BS / kWave;
  constexpr int U = 4;
  __shared__ double red[NW];
  __shared__ double red2[NW];
  int p, g = 0, G = 1;
  if (B.wg_prob) {
    p = B.wg_prob[blockIdx.x];
    g = B.wg_rank[blockIdx.x];
    G = B.groups[p];
  } else {
    p = B.order[blockIdx.x];
  }
  const int off = B.dof_off[p], na = B.n_active[p];
  const int ps0 = B.slice_off[p], ps1 = B.slice_off[p + 1];
  const int s0 = ps0 + (int)((long long)(ps1 - ps0) * g / G);
  const int s1 = ps0 + (int)((long long)(ps1 - ps0) * (g + 1) / G);
  unsigned epoch = 0;
  int nred = 0;
  // barrier of the problem's workgroups (a block barrier when G = 1)
  auto gbar = [&]() {
    __syncthreads();
    if (G == 1) return;
    ++epoch;
    if (threadIdx.x == 0) {
      __threadfence();
      atomicAdd(B.bar + p, 1u);
      const unsigned target = epoch * (unsigned)G;
      unsigned spins = 0;
      while (__hip_atomic_load(B.bar + p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++spins > kBatchSpinLimit) {
          atomicOr(B.bar_timeout, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      __threadfence();
    }
    __syncthreads();
  };
  double* x = B.x + off;
  const double* rhs = B.rhs + off;
  const float* dinv = B.dinv32 + off;
  float* ri = B.V + B.ws_off[p];
  float* rh = ri + na;
  float* pv = rh + na;
  float* v = pv + na;
  float* sv = v + na;
  float* tv = sv + na;
  float* dv = tv + na;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  // two sums at once (fixed order: waves, then the problem's workgroups)
  auto sum2 = [&](double a, double b, double& ra, double& rb) {
    a = group_sum<kWave>(a);
    b = group_sum<kWave>(b);
    if (lane == 0) { red[wv] = a; red2[wv] = b; }
    __syncthreads();
    double sa = 0, sb = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) { sa += red[i]; sb += red2[i]; }
    __syncthreads();
    if (G > 1) {
      // two slot buffers alternate: a workgroup can only overwrite a buffer
      // after the next barrier, which every reader of it has passed
      const int buf = (nred++) & 1;
      double* slots = B.part + ((size_t)p * 2 + buf) * kBatchMaxGroups * 2;
      if (threadIdx.x == 0) {
        slots[2 * g] = sa;
        slots[2 * g + 1] = sb;
      }
      gbar();
      sa = 0;
      sb = 0;
      for (int q = 0; q < G; ++q) {
        sa += __hip_atomic_load(slots + 2 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sb += __hip_atomic_load(slots + 2 * q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    ra = sa;
    rb = sb;
  };
  double rn0 = 0, rn = 0;
  int outer = 0, inner_total = 0;
  for (;;) {
    // FP64 residual; the inner right-hand side ri = rh = (float) r, d = 0
    double part = 0;
    wave_spmv<U, double, double>(B, B.S64, s0, s1, wv, NW, lane, [&](int c) { return x[c]; },
                                 [&](int lr, double acc) {
                                   const double d = rhs[lr] - acc;
                                   part += d * d;
                                   ri[lr] = (float)d;
                                   rh[lr] = (float)d;
                                   pv[lr] = 0.0f;
                                   v[lr] = 0.0f;
                                   dv[lr] = 0.0f;
                                 });
    double rr, dummy;
    sum2(part, 0.0, rr, dummy);
    rn = sqrt(rr);
    if (outer == 0) rn0 = rn;
    if (rn <= P.red * rn0 || rn <= P.tol || outer >= P.max_outer) break;
    ++outer;
    // rho = (rh, ri) of the FP32 copy
    double rho1 = 0;
    {
      double a = 0;
      wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) { const double q = ri[lr]; a += q * q; });
      sum2(a, 0.0, rho1, dummy);
    }
    const double stop = P.inner_red * sqrt(rho1);
    double rho = 1, alpha = 1, omega = 1;
    int it = 0;
    for (; it < P.max_inner; ++it) {
      if (!(fabs(rho1) > 0)) break;
      const float beta = (float)((rho1 / rho) * (alpha / omega));
      const float om = (float)omega;
      // p = ri + beta (p - omega v)
      wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) { pv[lr] = ri[lr] + beta * (pv[lr] - om * v[lr]); });
      gbar();
      // v = S D^-1 p ; (rh, v)
      double a = 0;
      wave_spmv<U, float, float>(B, B.S32, s0, s1, wv, NW, lane,
                                 [&](int c) { return (c < na) ? dinv[c] * pv[c] : 0.0f; },
                                 [&](int lr, float acc) { v[lr] = acc; a += (double)rh[lr] * (double)acc; });
      double rhv;
      sum2(a, 0.0, rhv, dummy);
      if (!(fabs(rhv) > 0)) break;
      alpha = rho1 / rhv;
      const float al = (float)alpha;
      // s = ri - alpha v ; ||s||
      a = 0;
      wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) {
        const float q = ri[lr] - al * v[lr];
        sv[lr] = q;
        a += (double)q * (double)q;
      });
      double ss;
      sum2(a, 0.0, ss, dummy);
      if (sqrt(ss) <= stop) {
        wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) { dv[lr] += al * dinv[lr] * pv[lr]; });
        ++it;
        break;
      }
      // t = S D^-1 s ; (t, s), (t, t)
      a = 0;
      double bb = 0;
      wave_spmv<U, float, float>(B, B.S32, s0, s1, wv, NW, lane,
                                 [&](int c) { return (c < na) ? dinv[c] * sv[c] : 0.0f; },
                                 [&](int lr, float acc) {
                                   tv[lr] = acc;
                                   a += (double)acc * (double)sv[lr];
                                   bb += (double)acc * (double)acc;
                                 });
      double ts, tt;
      sum2(a, bb, ts, tt);
      omega = (tt > 0) ? ts / tt : 0.0;
      const float om2 = (float)omega;
      // d += alpha D^-1 p + omega D^-1 s ; ri = s - omega t ; (rh, ri), ||ri||
      a = 0;
      bb = 0;
      wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) {
        dv[lr] += dinv[lr] * (al * pv[lr] + om2 * sv[lr]);
        const float q = sv[lr] - om2 * tv[lr];
        ri[lr] = q;
        a += (double)rh[lr] * (double)q;
        bb += (double)q * (double)q;
      });
      rho = rho1;
      double rr2;
      sum2(a, bb, rho1, rr2);
      if (sqrt(rr2) <= stop || !(fabs(omega) > 0)) { ++it; break; }
    }
    inner_total += it;
    // x += d (FP64)
    wave_rows<U>(B, s0, s1, wv, NW, lane, [&](int lr) { x[lr] += (double)dv[lr]; });
    gbar();
  }
  if (threadIdx.x == 0 && g == 0) {
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

// zero the per-problem barrier counters (a kernel node: the step is replayed
// from a graph)
__global__ void k_fill_u32(int n, unsigned int* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0u;
}

}  // namespace

int batch_solve_capacity(int block, int method) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  hipError_t e;
  if (method == 1) {
    switch (block) {
      case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_solve_bicgstab<256>, 256, 0); break;
      case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_solve_bicgstab<512>, 512, 0); break;
      default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_solve_bicgstab<1024>, 1024, 0); break;
    }
  } else {
    switch (block) {
      case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_solve<256>, 256, 0); break;
      case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_solve<512>, 512, 0); break;
      default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_solve<1024>, 1024, 0); break;
    }
  }
  return e == hipSuccess ? per * cus : 0;
}

void batch_solve(const BatchDev& B, const BatchSolveParams& P, int block, hipStream_t st) {
  if (P.method == 1) {
    const int grid = B.wg_prob ? B.n_wg : B.n_prob;
    if (B.wg_prob) {
      if (!B.groups || !B.bar || !B.part || !B.bar_timeout || B.n_wg < B.n_prob)
        throw std::runtime_error("batch: incomplete workgroup assignment");
      hipLaunchKernelGGL(k_fill_u32, dim3((B.n_prob + 255) / 256), dim3(256), 0, st, B.n_prob, B.bar);
    }
    switch (block) {
      case 256: hipLaunchKernelGGL((k_batch_solve_bicgstab<256>), dim3(grid), dim3(256), 0, st, B, P); break;
      case 512: hipLaunchKernelGGL((k_batch_solve_bicgstab<512>), dim3(grid), dim3(512), 0, st, B, P); break;
      default: hipLaunchKernelGGL((k_batch_solve_bicgstab<1024>), dim3(grid), dim3(1024), 0, st, B, P); break;
    }
    return;
  }
  if (P.m > kMaxRestart) throw std::runtime_error("batch: GMRES restart above 64");
  const int grid = B.wg_prob ? B.n_wg : B.n_prob;
  if (B.wg_prob) {
    if (!B.groups || !B.bar || !B.part || !B.bar_timeout || B.n_wg < B.n_prob)
      throw std::runtime_error("batch: incomplete workgroup assignment");
    hipLaunchKernelGGL(k_fill_u32, dim3((B.n_prob + 255) / 256), dim3(256), 0, st, B.n_prob, B.bar);
  }
  switch (block) {
    case 256: hipLaunchKernelGGL((k_batch_solve<256>), dim3(grid), dim3(256), 0, st, B, P); break;
    case 512: hipLaunchKernelGGL((k_batch_solve<512>), dim3(grid), dim3(512), 0, st, B, P); break;
    default: hipLaunchKernelGGL((k_batch_solve<1024>), dim3(grid), dim3(1024), 0, st, B, P); break;
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
