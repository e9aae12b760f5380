// Sparse/dense linear-algebra kernels of the solve path (gfx950).
//
//  * FE-matrix SpMV with the reference's Dirichlet semantics:
//    BlockFEMatrix::apply_scaled_add (src/Matrix/BlockFEMatrix.C:863-924) =
//    FEMatrix::multiplyActive on rows < ActiveBound (src/Matrix/FEMatrix.C:186-199)
//    plus y += a*x on the non-active rows (BlockVector::addScaledNonActive,
//    src/Matrix/BlockVector.C:212-228).
//  * fused defect d = b - A x with ||d||^2 (MultigridLevel::calculate_defect,
//    src/Multigrid/MultigridLevel.C:131-150).
//  * fused damped Jacobi step x' = x + w D^-1 (b - A x)
//    (JacobiSmoother::smooth, src/Multigrid/JacobiSmoother.C:22-30).
//  * plain CSR SpMV for the grid-transfer matrices.
//  * BLAS-1 with deterministic two-stage reductions (Ddot/Dnorm,
//    src/FE/Blas1.C:14-117 sum sequentially; here a fixed tree is used).
//
// Every kernel is HBM-bound: one row is handled by a group of LPR lanes of a
// 64-wide wavefront so that the row's values/columns are read coalesced.
#include <stdexcept>
#include <vector>
#include <type_traits>

#include "common.hpp"
#include "launch.hpp"

namespace pmx {

namespace {

enum SpmvMode { MODE_SCALED_ADD = 0, MODE_DEFECT = 1, MODE_JACOBI = 2, MODE_APPLY = 3 };

template <int LPR, int MODE, typename VT, typename XT>
__global__ __launch_bounds__(256) void k_fe_spmv(
    int n, int active, const int* __restrict__ rp, const int* __restrict__ ci,
    const VT* __restrict__ va, const float* __restrict__ rscale,
    const XT* __restrict__ x, const XT* __restrict__ b, const XT* __restrict__ dinv,
    XT* __restrict__ y, double a, double* __restrict__ partial) {
  __shared__ double red[16];
  const int lane = threadIdx.x & (LPR - 1);
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  using AT = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  AT acc = 0;
  double d2 = 0;
  if (row < n) {
    if (row < active) {
      const int s = rp[row], e = rp[row + 1];
      auto val = [&](int k) -> AT {
        if constexpr (sizeof(VT) == 2) return (AT)bf16_to_float(va[k]);
        else return (AT)va[k];
      };
      int k = s + lane;
      // four entries per lane loaded before their products are summed (in
      // the same order as one at a time)
      for (; k + 3 * LPR < e; k += 4 * LPR) {
        int c[4];
        AT v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = __builtin_nontemporal_load(ci + k + u * LPR);
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = val(k + u * LPR);
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += v[u] * (AT)x[c[u]];
      }
      for (; k < e; k += LPR) acc += val(k) * (AT)x[ci[k]];
      acc = group_sum<LPR>(acc);
      if (rscale) acc *= (AT)rscale[row];
    }
    if (lane == 0) {
      if (MODE == MODE_SCALED_ADD) {
        if (row < active) y[row] += (XT)(a * (double)acc);
        else y[row] += (XT)(a * (double)x[row]);
      } else if (MODE == MODE_APPLY) {
        y[row] = (row < active) ? (XT)acc : x[row];
      } else if (MODE == MODE_DEFECT) {
        const XT d = (row < active) ? (XT)((AT)b[row] - acc) : (XT)(b[row] - x[row]);
        y[row] = d;
        d2 = (double)d * (double)d;
      } else {  // MODE_JACOBI: y = x + a * dinv * (b - A x)
        const AT r = (row < active) ? ((AT)b[row] - acc) : (AT)(b[row] - x[row]);
        y[row] = (XT)((AT)x[row] + (AT)a * (AT)dinv[row] * r);
      }
    }
  }
  if (MODE == MODE_DEFECT) {
    const double s = block_sum(d2, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

template <int LPR, int MODE, typename VT, typename XT>
void launch_fe_spmv(const DevCSR& A, const VT* va, const float* rscale, const XT* x, const XT* b,
                    const XT* dinv, XT* y, double a, double* partial, hipStream_t st) {
  const int rows_per_block = 256 / LPR;
  const int grid = (A.n + rows_per_block - 1) / rows_per_block;
  hipLaunchKernelGGL((k_fe_spmv<LPR, MODE, VT, XT>), dim3(grid), dim3(256), 0, st, A.n, A.active,
                     A.row_ptr, A.col, va, rscale, x, b, dinv, y, a, partial);
}


// SELL-64-sigma FE-matrix kernel: one wavefront per slice of 64 rows, lane l
// owns row row[slice*64+l] and sums its entries sequentially in column order
// (the summation order of TMatrix::multiply).  Eight entries are loaded per
// step before their x values are gathered, to keep enough bytes in flight.
// Blocks beyond the active slices handle the non-active (identity) rows.
// FOLD (16-bit layout only): the row's Dirichlet-column products are not
// summed per launch; gsum[r] holds their sum (in column order, computed when
// the solution's Dirichlet part was last set) and starts the row sum, so the
// row is summed Dirichlet columns first, then the active columns in order.
// PACK 4 (folded layout only): slot j of lane l sits at s0 + (j/4)*256 +
// 4l + j%4, so a lane reads four offsets (8 bytes) and four values (8 bytes
// bf16, 32 bytes FP64) of its row per load instruction; the summation order
// is unchanged.
typedef double v2d_sell __attribute__((ext_vector_type(2)));
template <int MODE, typename VT, typename XT, bool REL16, bool FOLD = false, int PACK = 1>
__global__ __launch_bounds__(256) void k_sell_spmv(
    int n_slices, int sell_blocks, int n, int active, const int* __restrict__ sptr,
    const int* __restrict__ rowmap, const int* __restrict__ ci, const int16_t* __restrict__ off,
    const int* __restrict__ bd_ptr, const int* __restrict__ bd_col, const VT* __restrict__ bd_val,
    const VT* __restrict__ va, const float* __restrict__ rscale, const XT* __restrict__ x,
    const XT* __restrict__ b, const XT* __restrict__ dinv, XT* __restrict__ y, double a,
    double* __restrict__ partial, const XT* __restrict__ gsum = nullptr) {
  __shared__ double red[16];
  using AT = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  double d2 = 0;
  if ((int)blockIdx.x < sell_blocks) {
    // workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of slices so that its L2 sees neighbouring rows (and
    // the x entries they share).  sell_blocks is a multiple of 8.
    const int per = sell_blocks >> 3;
    const int vb = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    const int lane = threadIdx.x & 63;
    const int slice = vb * 4 + (int)(threadIdx.x >> 6);
    if (slice < n_slices) {
      const int r = rowmap[slice * 64 + lane];
      const int s0 = sptr[slice];
      const int width = (sptr[slice + 1] - s0) >> 6;
      const int* cp = ci + s0 + lane;
      const int16_t* op = off + s0 + lane;
      const int rb = r >= 0 ? r : 0;   // 16-bit layout: column = row + offset (empty lanes: offset 0)
      const VT* vp = va + s0 + lane;
      AT acc = 0;
      if constexpr (FOLD) acc = (AT)gsum[rb];
      int j = 0;
      if constexpr (PACK == 4) {
        const uint64_t* o4 = (const uint64_t*)(off + s0) + lane;
        auto unpack = [&](uint64_t o, int* c) {
#pragma unroll
          for (int t = 0; t < 4; ++t) c[t] = rb + (int)(int16_t)(uint16_t)(o >> (16 * t));
        };
        auto values = [&](int q, AT* v) {   // the four values of group q (slots 4q .. 4q+3)
          if constexpr (sizeof(VT) == 2) {
            const uint64_t w = __builtin_nontemporal_load((const uint64_t*)(va + s0) + (size_t)q * 64 + lane);
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] = (AT)bf16_to_float((uint16_t)(w >> (16 * t)));
          } else {
            const v2d_sell* p2 = (const v2d_sell*)(va + s0 + (size_t)q * 256 + 4 * lane);
            const v2d_sell a0 = __builtin_nontemporal_load(p2), a1 = __builtin_nontemporal_load(p2 + 1);
            v[0] = (AT)a0.x; v[1] = (AT)a0.y; v[2] = (AT)a1.x; v[3] = (AT)a1.y;
          }
        };
        for (; j + 8 <= width; j += 8) {
          const int q = j >> 2;
          int c[8];
          AT v[8];
          unpack(__builtin_nontemporal_load(o4 + (size_t)q * 64), c);
          unpack(__builtin_nontemporal_load(o4 + (size_t)(q + 1) * 64), c + 4);
          values(q, v);
          values(q + 1, v + 4);
#pragma unroll
          for (int u = 0; u < 8; ++u) acc += v[u] * (AT)x[c[u]];
        }
        for (; j < width; j += 4) {
          const int q = j >> 2;
          int c[4];
          AT v[4];
          unpack(__builtin_nontemporal_load(o4 + (size_t)q * 64), c);
          values(q, v);
#pragma unroll
          for (int u = 0; u < 4; ++u) acc += v[u] * (AT)x[c[u]];
        }
      }
      // the folded bf16 sweep keeps sixteen entries in flight per lane
      constexpr bool wide = PACK == 1 && FOLD;
      for (; wide && j + 16 <= width; j += 16) {
        int c[16];
        AT v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          if constexpr (REL16) c[u] = rb + (int)__builtin_nontemporal_load(op + (j + u) * 64);
          else c[u] = __builtin_nontemporal_load(cp + (j + u) * 64);
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          if constexpr (sizeof(VT) == 2) v[u] = (AT)bf16_to_float(__builtin_nontemporal_load(vp + (j + u) * 64));
          else v[u] = (AT)__builtin_nontemporal_load(vp + (j + u) * 64);
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) acc += v[u] * (AT)x[c[u]];
      }
      for (; PACK == 1 && j + 8 <= width; j += 8) {
        int c[8];
        AT v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if constexpr (REL16) c[u] = rb + (int)__builtin_nontemporal_load(op + (j + u) * 64);
          else c[u] = __builtin_nontemporal_load(cp + (j + u) * 64);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if constexpr (sizeof(VT) == 2) v[u] = (AT)bf16_to_float(__builtin_nontemporal_load(vp + (j + u) * 64));
          else v[u] = (AT)__builtin_nontemporal_load(vp + (j + u) * 64);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u] * (AT)x[c[u]];
      }
      for (; PACK == 1 && j < width; ++j) {
        int c;
        if constexpr (REL16) c = rb + (int)__builtin_nontemporal_load(op + j * 64);
        else c = __builtin_nontemporal_load(cp + j * 64);
        AT v;
        if constexpr (sizeof(VT) == 2) v = (AT)bf16_to_float(__builtin_nontemporal_load(vp + j * 64));
        else v = (AT)__builtin_nontemporal_load(vp + j * 64);
        acc += v * (AT)x[c];
      }
      if constexpr (REL16 && !FOLD) {
        // the Dirichlet-column entries, last in the row's column order
        if (r >= 0) {
          const int e = bd_ptr[r + 1];
          for (int k = bd_ptr[r]; k < e; ++k) {
            AT v;
            if constexpr (sizeof(VT) == 2) v = (AT)bf16_to_float(bd_val[k]);
            else v = (AT)bd_val[k];
            acc += v * (AT)x[bd_col[k]];
          }
        }
      }
      if (r >= 0) {
        if (rscale) acc *= (AT)rscale[r];
        if (MODE == MODE_SCALED_ADD) {
          y[r] += (XT)(a * (double)acc);
        } else if (MODE == MODE_APPLY) {
          y[r] = (XT)acc;
        } else if (MODE == MODE_DEFECT) {
          const XT d = (XT)((AT)b[r] - acc);
          y[r] = d;
          d2 = (double)d * (double)d;
        } else {
          const AT res = (AT)b[r] - acc;
          y[r] = (XT)((AT)x[r] + (AT)a * (AT)dinv[r] * res);
        }
      }
    }
  } else {
    const int i = active + (int)(blockIdx.x - sell_blocks) * blockDim.x + threadIdx.x;
    if (i < n) {
      if (MODE == MODE_SCALED_ADD) {
        y[i] += (XT)(a * (double)x[i]);
      } else if (MODE == MODE_APPLY) {
        y[i] = x[i];
      } else if (MODE == MODE_DEFECT) {
        const XT d = (XT)(b[i] - x[i]);
        y[i] = d;
        d2 = (double)d * (double)d;
      } else {
        const AT res = (AT)(b[i] - x[i]);
        y[i] = (XT)((AT)x[i] + (AT)a * (AT)dinv[i] * res);
      }
    }
  }
  if (MODE == MODE_DEFECT) {
    const double s = block_sum(d2, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// The pattern layout of the folded 16-bit SELL copy: the 64 rows of a slice
// share their active-column offsets (SellPattern::slice_pat), so a wave reads
// slot j's offset once (a uniform, scalar load) and streams only the values:
// 2 bytes per entry in bf16 instead of 4.  Row sums run in the order of
// k_sell_spmv<.., REL16, FOLD> (Dirichlet-column sum first, then the active
// columns), bitwise the same.  Empty lanes (the last slice of a pattern) use
// the slice's first row as base, so every x read stays inside the active
// rows; their values are 0 and their results are not stored.  GSUM false:
// the Dirichlet-column sums are taken as 0, valid when x vanishes on the
// Dirichlet rows (the corrections of an iterative refinement).
typedef float v4f_pat __attribute__((ext_vector_type(4)));
template <int MODE, typename VT, typename XT, bool GSUM, int PACK = 1, int WPB = 4, int CH = 16>
__global__ __launch_bounds__(64 * WPB) void k_sell_pat(
    int n_slices, int sell_blocks, int n, int active, const int* __restrict__ sptr,
    const int* __restrict__ rowmap, const int* __restrict__ slice_pat, const int* __restrict__ pat_off,
    const VT* __restrict__ va, const float* __restrict__ rscale, const XT* __restrict__ x,
    const XT* __restrict__ b, const XT* __restrict__ dinv, XT* __restrict__ y, double a,
    double* __restrict__ partial, const XT* __restrict__ gsum) {
  __shared__ double red[16];
  using AT = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  double d2 = 0;
  if ((int)blockIdx.x < sell_blocks) {
    const int per = sell_blocks >> 3;
    const int vb = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    const int lane = threadIdx.x & 63;
    const int slice = __builtin_amdgcn_readfirstlane(vb * WPB + (int)(threadIdx.x >> 6));
    if (slice < n_slices) {
      const int r = rowmap[slice * 64 + lane];
      const int r0 = __builtin_amdgcn_readfirstlane(rowmap[slice * 64]);
      const int s0 = __builtin_amdgcn_readfirstlane(sptr[slice]);
      const int width = __builtin_amdgcn_readfirstlane((sptr[slice + 1] - s0) >> 6);
      const int* po = pat_off + __builtin_amdgcn_readfirstlane(slice_pat[slice]);
      const int rb = r >= 0 ? r : r0;
      const VT* vp = va + s0 + lane;
      AT acc = 0;
      if constexpr (GSUM) acc = (AT)gsum[rb];
      // the row's own data, loaded before the sum so that it overlaps it
      AT br = 0, dr = 0, xr = 0, sr = 1;
      if (r >= 0) {
        if constexpr (MODE == MODE_DEFECT || MODE == MODE_JACOBI) br = (AT)b[r];
        if constexpr (MODE == MODE_JACOBI) {
          dr = (AT)dinv[r];
          xr = (AT)x[r];
        }
        if (rscale) sr = (AT)rscale[r];
      }
      auto val = [&](int j) -> AT {
        if constexpr (sizeof(VT) == 2) return (AT)bf16_to_float(__builtin_nontemporal_load(vp + j * 64));
        else return (AT)__builtin_nontemporal_load(vp + j * 64);
      };
      int j = 0;
      if constexpr (PACK == 4) {
        // slot j of this lane at s0 + (j / 4) 256 + 4 lane + j % 4: one load
        // per four slots; widths are multiples of 4 (padding: offset 0, value 0)
        auto group = [&](int q, AT* v) {
          const VT* g = va + s0 + (size_t)q * 256 + 4 * lane;
          if constexpr (sizeof(VT) == 2) {
            const uint64_t w = __builtin_nontemporal_load((const uint64_t*)g);
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t] = (AT)bf16_to_float((uint16_t)(w >> (16 * t)));
          } else if constexpr (sizeof(VT) == 4) {
            const v4f_pat w = __builtin_nontemporal_load((const v4f_pat*)g);
            v[0] = (AT)w.x; v[1] = (AT)w.y; v[2] = (AT)w.z; v[3] = (AT)w.w;
          } else {
            const v2d_sell a0 = __builtin_nontemporal_load((const v2d_sell*)g);
            const v2d_sell a1 = __builtin_nontemporal_load((const v2d_sell*)g + 1);
            v[0] = (AT)a0.x; v[1] = (AT)a0.y; v[2] = (AT)a1.x; v[3] = (AT)a1.y;
          }
        };
        for (; j + 16 <= width; j += 16) {
          int c[16];
          AT v[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) c[u] = rb + po[j + u];
#pragma unroll
          for (int t = 0; t < 4; ++t) group((j >> 2) + t, v + 4 * t);
#pragma unroll
          for (int u = 0; u < 16; ++u) acc += v[u] * (AT)x[c[u]];
        }
        for (; j < width; j += 4) {
          int c[4];
          AT v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) c[u] = rb + po[j + u];
          group(j >> 2, v);
#pragma unroll
          for (int u = 0; u < 4; ++u) acc += v[u] * (AT)x[c[u]];
        }
      }
      for (; PACK == 1 && j + CH <= width; j += CH) {
        int c[CH];
        AT v[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) c[u] = rb + po[j + u];
#pragma unroll
        for (int u = 0; u < CH; ++u) v[u] = val(j + u);
#pragma unroll
        for (int u = 0; u < CH; ++u) acc += v[u] * (AT)x[c[u]];
      }
      for (; PACK == 1 && CH > 16 && j + 16 <= width; j += 16) {
        int c[16];
        AT v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) c[u] = rb + po[j + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = val(j + u);
#pragma unroll
        for (int u = 0; u < 16; ++u) acc += v[u] * (AT)x[c[u]];
      }
      for (; PACK == 1 && j + 8 <= width; j += 8) {
        int c[8];
        AT v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = rb + po[j + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = val(j + u);
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u] * (AT)x[c[u]];
      }
      for (; PACK == 1 && j < width; ++j) acc += val(j) * (AT)x[rb + po[j]];
      if (r >= 0) {
        if (rscale) acc *= sr;
        if (MODE == MODE_SCALED_ADD) {
          y[r] += (XT)(a * (double)acc);
        } else if (MODE == MODE_APPLY) {
          y[r] = (XT)acc;
        } else if (MODE == MODE_DEFECT) {
          const XT d = (XT)(br - acc);
          y[r] = d;
          d2 = (double)d * (double)d;
        } else {
          const AT res = br - acc;
          y[r] = (XT)(xr + (AT)a * dr * res);
        }
      }
    }
  } else {
    const int i = active + (int)(blockIdx.x - sell_blocks) * blockDim.x + threadIdx.x;
    if (i < n) {
      if (MODE == MODE_SCALED_ADD) {
        y[i] += (XT)(a * (double)x[i]);
      } else if (MODE == MODE_APPLY) {
        y[i] = x[i];
      } else if (MODE == MODE_DEFECT) {
        const XT d = (XT)(b[i] - x[i]);
        y[i] = d;
        d2 = (double)d * (double)d;
      } else {
        const AT res = (AT)(b[i] - x[i]);
        y[i] = (XT)((AT)x[i] + (AT)a * (AT)dinv[i] * res);
      }
    }
  }
  if (MODE == MODE_DEFECT) {
    const double s = block_sum(d2, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// y_k += a A x_k for three vectors at once (x_k = x + k * ld, y_k likewise):
// the velocity block of an NSTYPE-1 saddle-point operator acts on all three
// components, so its SELL slots are read once for the three products.  Each
// row sum runs in the order of k_sell_spmv (bitwise the same results).
__global__ __launch_bounds__(256) void k_sell_spmv3(int n_slices, int sell_blocks, int n, int active, size_t ld,
                                                    const int* __restrict__ sptr, const int* __restrict__ rowmap,
                                                    const int* __restrict__ ci, const double* __restrict__ va,
                                                    const double* __restrict__ x, double* __restrict__ y, double a) {
  if ((int)blockIdx.x < sell_blocks) {
    const int per = sell_blocks >> 3;
    const int vb = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    const int lane = threadIdx.x & 63;
    const int slice = vb * 4 + (int)(threadIdx.x >> 6);
    if (slice >= n_slices) return;
    const int r = rowmap[slice * 64 + lane];
    const int s0 = sptr[slice];
    const int width = (sptr[slice + 1] - s0) >> 6;
    const int* cp = ci + s0 + lane;
    const double* vp = va + s0 + lane;
    double a0 = 0, a1 = 0, a2 = 0;
    int j = 0;
    for (; j + 4 <= width; j += 4) {
      int c[4];
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = __builtin_nontemporal_load(cp + (j + u) * 64);
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(vp + (j + u) * 64);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a0 += v[u] * x[c[u]];
        a1 += v[u] * x[ld + c[u]];
        a2 += v[u] * x[2 * ld + c[u]];
      }
    }
    for (; j < width; ++j) {
      const int c = __builtin_nontemporal_load(cp + j * 64);
      const double v = __builtin_nontemporal_load(vp + j * 64);
      a0 += v * x[c];
      a1 += v * x[ld + c];
      a2 += v * x[2 * ld + c];
    }
    if (r >= 0) {
      y[r] += a * a0;
      y[ld + r] += a * a1;
      y[2 * ld + r] += a * a2;
    }
  } else {
    const int i = active + (int)(blockIdx.x - sell_blocks) * blockDim.x + threadIdx.x;
    if (i < n) {
      y[i] += a * x[i];
      y[ld + i] += a * x[ld + i];
      y[2 * ld + i] += a * x[2 * ld + i];
    }
  }
}

// MODE_JACOBI of k_sell_spmv for an x that vanishes on the active rows: the
// products with those zeros drop out of each row sum exactly (a*0 added to a
// partial sum leaves it unchanged), so only the row's Dirichlet-column
// entries are summed, in the same order and with the same values (a compact
// copy of those SELL slots), and the result equals the full sweep bit for
// bit.  One thread per row; reads b, D^-1 and writes x' (about 16 bytes per
// row in FP32) instead of streaming the whole matrix.
template <typename VT, typename XT>
__global__ __launch_bounds__(256) void k_sell_jacobi_from_zero(
    int n, int active, const int* __restrict__ bd_ptr, const int* __restrict__ bd_col,
    const VT* __restrict__ bd_val, const float* __restrict__ rscale,
    const XT* __restrict__ x, const XT* __restrict__ b, const XT* __restrict__ dinv,
    XT* __restrict__ y, double a, XT* __restrict__ gsum) {
  using AT = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i < active) {
    AT acc = 0;
    const int e = bd_ptr[i + 1];
    for (int k = bd_ptr[i]; k < e; ++k) {
      AT v;
      if constexpr (sizeof(VT) == 2) v = (AT)bf16_to_float(bd_val[k]);
      else v = (AT)bd_val[k];
      acc += v * (AT)x[bd_col[k]];
    }
    if (gsum) gsum[i] = (XT)acc;   // the Dirichlet-column sum of the folded layout
    if (rscale) acc *= (AT)rscale[i];
    const AT res = (AT)b[i] - acc;
    y[i] = (XT)((AT)a * (AT)dinv[i] * res);
  } else {
    const AT res = (AT)(b[i] - x[i]);
    y[i] = (XT)((AT)x[i] + (AT)a * (AT)dinv[i] * res);
  }
}

// k_sell_jacobi_from_zero for an x that vanishes on every row (Dirichlet
// rows too): no Dirichlet-column products, and x is not read at all (so the
// caller need not clear it); y = w D^-1 b on every row, the expressions of
// the sweep with x = 0 (the Dirichlet rows: x + w D^-1 (b - x) = w D^-1 b)
template <typename XT>
__global__ __launch_bounds__(256) void k_jacobi_from_zero_nodir(int n, const XT* __restrict__ b,
                                                                const XT* __restrict__ dinv, XT* __restrict__ y,
                                                                double a) {
  using AT = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const AT res = (AT)b[i];
  y[i] = (XT)((AT)a * (AT)dinv[i] * res);
}

// gsum[r] = sum of the Dirichlet-column products of active row r, in column
// order (the folded 16-bit layout; before the row scale)
template <typename VT, typename XT>
__global__ __launch_bounds__(256) void k_sell_dirichlet_sums(int active, const int* __restrict__ bd_ptr,
                                                             const int* __restrict__ bd_col,
                                                             const VT* __restrict__ bd_val,
                                                             const XT* __restrict__ x, XT* __restrict__ gsum) {
  using AT = typename std::conditional<sizeof(XT) == 8, double, float>::type;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= active) return;
  AT acc = 0;
  const int e = bd_ptr[i + 1];
  for (int k = bd_ptr[i]; k < e; ++k) {
    AT v;
    if constexpr (sizeof(VT) == 2) v = (AT)bf16_to_float(bd_val[k]);
    else v = (AT)bd_val[k];
    acc += v * (AT)x[bd_col[k]];
  }
  gsum[i] = (XT)acc;
}

// four slices (waves) per workgroup
int sell_grid(const DevCSR& A, int* sell_blocks) {
  const int w = 4;
  *sell_blocks = (((A.sell.n_slices + w - 1) / w + 7) / 8) * 8;
  return *sell_blocks + (A.n - A.active + 64 * w - 1) / (64 * w);
}

template <int MODE, typename VT, typename XT>
void launch_sell(const DevCSR& A, const VT* va, const VT* bd_val, const float* rscale, const XT* x,
                 const XT* b, const XT* dinv, XT* y, double a, double* partial, hipStream_t st) {
  int sb = 0;
  const int grid = sell_grid(A, &sb);
  if (A.sell.off && A.sell.n_bd > 0 && (!bd_val || !A.sell.bd_ptr || !A.sell.bd_col))
    throw std::runtime_error("16-bit SELL layout without its Dirichlet-column values");
  if (A.sell.n_slices > 0 && (A.sell.total == 0 || !va || (!A.sell.off && !A.sell.col)))
    throw std::runtime_error("SELL copy without slots, values or columns");
  if (A.sell.pack != 1 && !(A.sell.pack == 4 && A.sell.off && A.sell.gsum))
    throw std::runtime_error("packed SELL slots need the folded 16-bit layout");
  if (A.sell.slice_pat && A.sell.off && A.sell.gsum) {
    if (A.sell.gsum_bytes != (int)sizeof(XT)) throw std::runtime_error("folded SELL layout: sums of another precision");
    const XT* g = A.sell.dir_zero ? nullptr : (const XT*)A.sell.gsum;
#define PMX_PAT(GS, PK, C)                                                                                           \
  hipLaunchKernelGGL((k_sell_pat<MODE, VT, XT, GS, PK, 4, C>), dim3(grid), dim3(256), 0, st, A.sell.n_slices, sb,  \
                     A.n, A.active, A.sell.slice_ptr, A.sell.row, A.sell.slice_pat, A.sell.pat_off, va, rscale, x,  \
                     b, dinv, y, a, partial, g)
    if (A.sell.pack == 4) {
      if (g) PMX_PAT(true, 4, 16); else PMX_PAT(false, 4, 16);
    } else {
      if (g) PMX_PAT(true, 1, 16); else PMX_PAT(false, 1, 16);
    }
#undef PMX_PAT
  } else if (A.sell.pack == 4) {
    if (A.sell.gsum_bytes != (int)sizeof(XT)) throw std::runtime_error("folded SELL layout: sums of another precision");
    hipLaunchKernelGGL((k_sell_spmv<MODE, VT, XT, true, true, 4>), dim3(grid), dim3(256), 0, st, A.sell.n_slices,
                       sb, A.n, A.active, A.sell.slice_ptr, A.sell.row, A.sell.col, A.sell.off, A.sell.bd_ptr,
                       A.sell.bd_col, bd_val, va, rscale, x, b, dinv, y, a, partial, (const XT*)A.sell.gsum);
  } else if (A.sell.off && A.sell.gsum) {
    if (A.sell.gsum_bytes != (int)sizeof(XT)) throw std::runtime_error("folded SELL layout: sums of another precision");
    hipLaunchKernelGGL((k_sell_spmv<MODE, VT, XT, true, true>), dim3(grid), dim3(256), 0, st, A.sell.n_slices, sb,
                       A.n, A.active, A.sell.slice_ptr, A.sell.row, A.sell.col, A.sell.off, A.sell.bd_ptr,
                       A.sell.bd_col, bd_val, va, rscale, x, b, dinv, y, a, partial,
                       (const XT*)A.sell.gsum);
  } else if (A.sell.off)
    hipLaunchKernelGGL((k_sell_spmv<MODE, VT, XT, true>), dim3(grid), dim3(256), 0, st, A.sell.n_slices, sb,
                       A.n, A.active, A.sell.slice_ptr, A.sell.row, A.sell.col, A.sell.off, A.sell.bd_ptr,
                       A.sell.bd_col, bd_val, va, rscale, x, b, dinv, y, a, partial);
  else
    hipLaunchKernelGGL((k_sell_spmv<MODE, VT, XT, false>), dim3(grid), dim3(256), 0, st, A.sell.n_slices, sb,
                       A.n, A.active, A.sell.slice_ptr, A.sell.row, A.sell.col, A.sell.off, A.sell.bd_ptr,
                       A.sell.bd_col, bd_val, va, rscale, x, b, dinv, y, a, partial);
}

// SELL values from the CSR values through the slot -> position map; each
// thread keeps 16 slots in flight (index loads, then the gathers, then the
// stores) instead of one dependent chain at a time
template <typename T>
__global__ void k_sell_fill(size_t total, const int* __restrict__ src, const double* __restrict__ v,
                            T* __restrict__ out) {
  constexpr int U = 16;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; k + (U - 1) * stride < total; k += U * stride) {
    int s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = src[k + u * stride];
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = (s[u] >= 0) ? v[s[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) out[k + u * stride] = (T)x[u];
  }
  for (; k < total; k += stride) {
    const int s1 = src[k];
    out[k] = (T)((s1 >= 0) ? v[s1] : 0.0);
  }
}

// one wavefront per slice: lane l converts the entries of its row
__global__ __launch_bounds__(256) void k_sell_fill_bf16(int n_slices, const int* __restrict__ sptr,
                                                        const int* __restrict__ rowmap,
                                                        const int* __restrict__ src,
                                                        const double* __restrict__ v,
                                                        const float* __restrict__ rscale,
                                                        uint16_t* __restrict__ out, int pack) {
  const int lane = threadIdx.x & 63;
  const int slice = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (slice >= n_slices) return;
  const int r = rowmap[slice * 64 + lane];
  const double inv = (r >= 0) ? 1.0 / (double)rscale[r] : 0.0;
  const int s0 = sptr[slice], width = (sptr[slice + 1] - s0) >> 6;
  auto slot = [&](int j) { return s0 + (j / pack) * 64 * pack + lane * pack + j % pack; };
  auto conv = [&](int s) -> uint16_t {
    const float f = (s >= 0) ? (float)(v[s] * inv) : 0.0f;
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  };
  // 16 slots in flight: the index loads, then the gathers, then the stores
  int j = 0;
  for (; j + 16 <= width; j += 16) {
    int s[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) s[u] = src[slot(j + u)];
    uint16_t h[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) h[u] = conv(s[u]);
#pragma unroll
    for (int u = 0; u < 16; ++u) out[slot(j + u)] = h[u];
  }
  for (; j < width; ++j) out[slot(j)] = conv(src[slot(j)]);
}

template <int MODE, typename VT, typename XT>
void dispatch_fe_spmv(const DevCSR& A, const VT* va, const float* rscale, const XT* x, const XT* b,
                      const XT* dinv, XT* y, double a, double* partial, hipStream_t st) {
  const bool have_sell = (sizeof(VT) == 2) ? (A.sell.val_bf16 != nullptr) : (A.sell.val != nullptr);
  if (have_sell && A.lpr == 0) {
    if constexpr (sizeof(VT) == 2)
      launch_sell<MODE, VT, XT>(A, (const VT*)A.sell.val_bf16, (const VT*)A.sell.bd_val_bf16, rscale, x, b,
                                dinv, y, a, partial, st);
    else
      launch_sell<MODE, VT, XT>(A, (const VT*)A.sell.val, (const VT*)A.sell.bd_val, rscale, x, b, dinv, y, a,
                                partial, st);
    return;
  }
  switch (A.lpr) {
    case 4: launch_fe_spmv<4, MODE>(A, va, rscale, x, b, dinv, y, a, partial, st); break;
    case 8: launch_fe_spmv<8, MODE>(A, va, rscale, x, b, dinv, y, a, partial, st); break;
    case 16: launch_fe_spmv<16, MODE>(A, va, rscale, x, b, dinv, y, a, partial, st); break;
    case 32: launch_fe_spmv<32, MODE>(A, va, rscale, x, b, dinv, y, a, partial, st); break;
    default: launch_fe_spmv<64, MODE>(A, va, rscale, x, b, dinv, y, a, partial, st); break;
  }
}

// y = alpha * P x + beta * y over all rows (grid transfer)
// VT = uint16_t: the values are bf16 numbers that equal the FP64 ones
// exactly (nested-space interpolation weights such as 3/8 or -9/256), so the
// result is the same as with FP64 values, for 2 instead of 8 bytes each
template <int LPR, typename T, typename VT>
__global__ __launch_bounds__(256) void k_csr_spmv(int n, const int* __restrict__ rp,
                                                   const int* __restrict__ ci,
                                                   const VT* __restrict__ va,
                                                   const T* __restrict__ x, T* __restrict__ y,
                                                   double alpha, double beta, int row_end,
                                                   int zero_rest) {
  const int lane = threadIdx.x & (LPR - 1);
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  if (row >= n) return;
  if (row >= row_end) {
    if (zero_rest && lane == 0) y[row] = (T)0;
    return;
  }
  double acc = 0;
  const int s = rp[row], e = rp[row + 1];
  for (int k = s + lane; k < e; k += LPR) {
    double v;
    if constexpr (sizeof(VT) == 2) v = (double)bf16_to_float(va[k]);
    else v = va[k];
    acc += v * (double)x[ci[k]];
  }
  acc = group_sum<LPR>(acc);
  if (lane == 0) {
    double r = alpha * acc;
    if (beta != 0.0) r += beta * (double)y[row];
    y[row] = (T)r;
  }
}

// Grid transfer y = alpha P x + beta y on rows [0, n_sell) from the SELL copy
// of P (one lane per row, the row's entries in column order: the sequential
// CSR sum), rows [n_sell, n) set to zero (zero_rest) or left alone.  The
// slices are remapped across the XCDs as in k_sell_spmv.  Transfer rows are
// short and uneven (1 to 125 entries for Q2): sorting by length within
// windows gives slices of one length, with no lanes idling on short rows.
template <typename T, typename VT>
__global__ __launch_bounds__(256) void k_sell_transfer(int n_slices, int sell_blocks, int n, int n_sell,
                                                       const int* __restrict__ sptr,
                                                       const int* __restrict__ rowmap,
                                                       const int* __restrict__ ci, const VT* __restrict__ va,
                                                       const T* __restrict__ x, T* __restrict__ y,
                                                       double alpha, double beta, int zero_rest) {
  using AT = typename std::conditional<sizeof(T) == 8, double, float>::type;
  if ((int)blockIdx.x < sell_blocks) {
    const int per = sell_blocks >> 3;
    const int vb = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    const int lane = threadIdx.x & 63;
    const int slice = vb * 4 + (int)(threadIdx.x >> 6);
    if (slice >= n_slices) return;
    const int r = rowmap[slice * 64 + lane];
    const int s0 = sptr[slice];
    const int width = (sptr[slice + 1] - s0) >> 6;
    const int* cp = ci + s0 + lane;
    const VT* vp = va + s0 + lane;
    AT acc = 0;
    int j = 0;
    for (; j + 4 <= width; j += 4) {
      int c[4];
      AT v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) c[u] = __builtin_nontemporal_load(cp + (j + u) * 64);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if constexpr (sizeof(VT) == 2) v[u] = (AT)bf16_to_float(__builtin_nontemporal_load(vp + (j + u) * 64));
        else v[u] = (AT)__builtin_nontemporal_load(vp + (j + u) * 64);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc += v[u] * (AT)x[c[u]];
    }
    for (; j < width; ++j) {
      AT v;
      if constexpr (sizeof(VT) == 2) v = (AT)bf16_to_float(__builtin_nontemporal_load(vp + j * 64));
      else v = (AT)__builtin_nontemporal_load(vp + j * 64);
      acc += v * (AT)x[__builtin_nontemporal_load(cp + j * 64)];
    }
    if (r >= 0) {
      AT res = (AT)alpha * acc;
      if (beta != 0.0) res += (AT)beta * (AT)y[r];
      y[r] = (T)res;
    }
  } else if (zero_rest) {
    const int i = n_sell + (int)(blockIdx.x - sell_blocks) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (T)0;
  }
}

// y = alpha P x + beta y on the rows of the pattern slices of P: one wave per
// slice, the slice's column offsets and weights read once per wave (uniform
// scalar loads), each lane gathers x[base + offset]; the row sum runs in
// column order as in k_sell_transfer (the left-out Dirichlet columns meet
// zeros there, which leave a sum unchanged).
template <typename T>
__device__ __forceinline__ void pat_transfer_block(int block, int n_slices, int sell_blocks,
                                                   const int* __restrict__ rowmap, const int* __restrict__ base,
                                                   const int* __restrict__ slice_pat, const int* __restrict__ pat_ptr,
                                                   const int* __restrict__ pat_col, const double* __restrict__ pat_val,
                                                   const T* __restrict__ x, T* __restrict__ y, double alpha,
                                                   double beta) {
  using AT = typename std::conditional<sizeof(T) == 8, double, float>::type;
  const int per = sell_blocks >> 3;
  const int vb = (block & 7) * per + (block >> 3);
  const int lane = threadIdx.x & 63;
  const int slice = __builtin_amdgcn_readfirstlane(vb * 4 + (int)(threadIdx.x >> 6));
  if (slice >= n_slices) return;
  const int r = rowmap[slice * 64 + lane];
  const T* xb = x + base[slice * 64 + lane];
  const int p = __builtin_amdgcn_readfirstlane(slice_pat[slice]);
  const int j0 = __builtin_amdgcn_readfirstlane(pat_ptr[p]);
  const int j1 = __builtin_amdgcn_readfirstlane(pat_ptr[p + 1]);
  AT acc = 0;
  int j = j0;
  for (; j + 8 <= j1; j += 8) {
    AT xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = (AT)xb[pat_col[j + u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += (AT)pat_val[j + u] * xv[u];
  }
  for (; j < j1; ++j) acc += (AT)pat_val[j] * (AT)xb[pat_col[j]];
  if (r >= 0) {
    AT res = (AT)alpha * acc;
    if (beta != 0.0) res += (AT)beta * (AT)y[r];
    y[r] = (T)res;
  }
}

// One launch for a transfer with pattern slices: workgroups [0, pat_blocks)
// run the pattern slices, the next ones the SELL copy of the other rows and
// the zero tail (k_sell_transfer's layout of blocks, shifted).
template <typename T, typename VT>
__global__ __launch_bounds__(256) void k_pat_transfer(int pat_slices, int pat_blocks, const int* __restrict__ prow,
                                                      const int* __restrict__ pbase, const int* __restrict__ slice_pat,
                                                      const int* __restrict__ pat_ptr, const int* __restrict__ pat_col,
                                                      const double* __restrict__ pat_val, int n_slices,
                                                      int sell_blocks, int n, int n_sell,
                                                      const int* __restrict__ sptr, const int* __restrict__ rowmap,
                                                      const int* __restrict__ ci, const VT* __restrict__ va,
                                                      const T* __restrict__ x, T* __restrict__ y, double alpha,
                                                      double beta, int zero_rest) {
  if ((int)blockIdx.x < pat_blocks) {
    pat_transfer_block<T>((int)blockIdx.x, pat_slices, pat_blocks, prow, pbase, slice_pat, pat_ptr, pat_col, pat_val,
                          x, y, alpha, beta);
    return;
  }
  const int block = (int)blockIdx.x - pat_blocks;
  using AT = typename std::conditional<sizeof(T) == 8, double, float>::type;
  if (block < sell_blocks) {
    const int per = sell_blocks >> 3;
    const int vb = (block & 7) * per + (block >> 3);
    const int lane = threadIdx.x & 63;
    const int slice = vb * 4 + (int)(threadIdx.x >> 6);
    if (slice >= n_slices) return;
    const int r = rowmap[slice * 64 + lane];
    const int s0 = sptr[slice];
    const int width = (sptr[slice + 1] - s0) >> 6;
    AT acc = 0;
    for (int j = 0; j < width; ++j) {
      AT v;
      if constexpr (sizeof(VT) == 2) v = (AT)bf16_to_float(va[s0 + j * 64 + lane]);
      else v = (AT)va[s0 + j * 64 + lane];
      acc += v * (AT)x[ci[s0 + j * 64 + lane]];
    }
    if (r >= 0) {
      AT res = (AT)alpha * acc;
      if (beta != 0.0) res += (AT)beta * (AT)y[r];
      y[r] = (T)res;
    }
  } else if (zero_rest) {
    const int i = n_sell + (block - sell_blocks) * (int)blockDim.x + (int)threadIdx.x;
    if (i < n) y[i] = (T)0;
  }
}

// The same transfer for three vectors (x + k ldx -> y + k ldy, k = 0, 1, 2):
// the velocity components of a saddle-point level share one interpolation
// matrix, read once here; each component's row sum as in k_csr_spmv.
template <int LPR, typename VT>
__global__ __launch_bounds__(256) void k_csr_spmv3(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                    const VT* __restrict__ va, const double* __restrict__ x,
                                                    size_t ldx, double* __restrict__ y, size_t ldy, double alpha,
                                                    double beta, int row_end, int zero_rest) {
  const int lane = threadIdx.x & (LPR - 1);
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  if (row >= n) return;
  if (row >= row_end) {
    if (zero_rest && lane == 0) {
      y[row] = 0.0;
      y[ldy + row] = 0.0;
      y[2 * ldy + row] = 0.0;
    }
    return;
  }
  double a0 = 0, a1 = 0, a2 = 0;
  const int s = rp[row], e = rp[row + 1];
  for (int k = s + lane; k < e; k += LPR) {
    double v;
    if constexpr (sizeof(VT) == 2) v = (double)bf16_to_float(va[k]);
    else v = va[k];
    const int c = ci[k];
    a0 += v * x[c];
    a1 += v * x[ldx + c];
    a2 += v * x[2 * ldx + c];
  }
  a0 = group_sum<LPR>(a0);
  a1 = group_sum<LPR>(a1);
  a2 = group_sum<LPR>(a2);
  if (lane == 0) {
    double r0 = alpha * a0, r1 = alpha * a1, r2 = alpha * a2;
    if (beta != 0.0) {
      r0 += beta * y[row];
      r1 += beta * y[ldy + row];
      r2 += beta * y[2 * ldy + row];
    }
    y[row] = r0;
    y[ldy + row] = r1;
    y[2 * ldy + row] = r2;
  }
}

// Divergence blocks of an NSTYPE-1 saddle-point operator, all three at once
// (one pattern, values v1 v2 v3):
//   k_csr_b3:  y[r] += a (B1 x1 + B2 x2 + B3 x3)[r]   (x_k = x + k ldx)
//   k_csr_bt3: y_k[r] += a (B_k^T p)[r]               (y_k = y + k ldy)
// for rows r < row_end; the pattern and the gathered p are read once, and
// every product is summed and added to y in the order of three k_csr_spmv
// calls with the same lanes per row (bitwise the same; csr_b3 runs 16 lanes
// per pressure row, csr_bt3 8 per velocity row).
template <int LPR>
__global__ __launch_bounds__(256) void k_csr_b3(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                const double* __restrict__ v1, const double* __restrict__ v2,
                                                const double* __restrict__ v3, const double* __restrict__ x,
                                                size_t ldx, double* __restrict__ y, double a, int row_end) {
  const int lane = threadIdx.x & (LPR - 1);
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  if (row >= n || row >= row_end) return;
  double a0 = 0, a1 = 0, a2 = 0;
  const int s = rp[row], e = rp[row + 1];
  for (int k = s + lane; k < e; k += LPR) {
    const int c = ci[k];
    a0 += v1[k] * x[c];
    a1 += v2[k] * x[ldx + c];
    a2 += v3[k] * x[2 * ldx + c];
  }
  a0 = group_sum<LPR>(a0);
  a1 = group_sum<LPR>(a1);
  a2 = group_sum<LPR>(a2);
  if (lane == 0) {
    double t = a * a0 + y[row];
    t = a * a1 + t;
    y[row] = a * a2 + t;
  }
}
template <int LPR>
__global__ __launch_bounds__(256) void k_csr_bt3(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const double* __restrict__ v1, const double* __restrict__ v2,
                                                 const double* __restrict__ v3, const double* __restrict__ p,
                                                 double* __restrict__ y, size_t ldy, double a, int row_end) {
  const int lane = threadIdx.x & (LPR - 1);
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  if (row >= n || row >= row_end) return;
  double a0 = 0, a1 = 0, a2 = 0;
  const int s = rp[row], e = rp[row + 1];
  for (int k = s + lane; k < e; k += LPR) {
    const double pc = p[ci[k]];
    a0 += v1[k] * pc;
    a1 += v2[k] * pc;
    a2 += v3[k] * pc;
  }
  a0 = group_sum<LPR>(a0);
  a1 = group_sum<LPR>(a1);
  a2 = group_sum<LPR>(a2);
  if (lane == 0) {
    y[row] = a * a0 + y[row];
    y[ldy + row] = a * a1 + y[ldy + row];
    y[2 * ldy + row] = a * a2 + y[2 * ldy + row];
  }
}

// ---------------------------------------------------------------- BLAS-1
constexpr int kRedBlocks = 1024;

template <typename T>
__global__ __launch_bounds__(256) void k_dot_partial(int n, const T* __restrict__ x,
                                                      const T* __restrict__ y,
                                                      double* __restrict__ partial) {
  __shared__ double red[16];
  double s = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s += (double)x[i] * (double)y[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[0] = sum(partial[0..m)) (op 0), sqrt of it (op 1); optional scale out
__global__ __launch_bounds__(1024) void k_reduce_partials(int m, const double* __restrict__ partial,
                                                          double* __restrict__ out, int op) {
  __shared__ double red[16];
  double s = 0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) s += partial[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = (op == 1) ? sqrt(s) : s;
}

template <typename T>
__global__ void k_axpy(int n, double a, const double* __restrict__ a_dev, double a_dev_sign,
                       const T* __restrict__ x, T* __restrict__ y) {
  const double aa = a_dev ? a_dev_sign * a_dev[0] : a;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = (T)((double)y[i] + aa * (double)x[i]);
}

// r -= h[k] * v ; partial dot of the updated r with w (pipelined modified
// Gram-Schmidt step of Iteration_gmres::flexible_gmres, src/Solver/Iteration_gmres.C:452-466)
__global__ __launch_bounds__(256) void k_axpy_dot(int n, const double* __restrict__ h,
                                                  const double* __restrict__ v,
                                                  double* __restrict__ r,
                                                  const double* __restrict__ w,
                                                  double* __restrict__ partial) {
  __shared__ double red[16];
  const double hk = h ? h[0] : 0.0;
  double s = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double ri = r[i];
    if (h) { ri -= hk * v[i]; r[i] = ri; }
    s += ri * (w ? w[i] : ri);
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// the same step with the previous dot product still in per-block partial
// sums: every workgroup first sums partial_in (in one fixed order, so all of
// them get the same h) and workgroup 0 records h; this removes the separate
// reduction launch between two Gram-Schmidt steps
template <typename T>
__global__ __launch_bounds__(256) void k_axpy_dot_reduced(int n, int m_in, const double* __restrict__ partial_in,
                                                          double* __restrict__ h_out,
                                                          const T* __restrict__ v,
                                                          T* __restrict__ r,
                                                          const T* __restrict__ w,
                                                          double* __restrict__ partial) {
  __shared__ double red[16];
  __shared__ double hs;
  double p = 0;
  for (int i = threadIdx.x; i < m_in; i += blockDim.x) p += partial_in[i];
  p = block_sum(p, red);
  if (threadIdx.x == 0) {
    hs = p;
    if (blockIdx.x == 0) h_out[0] = p;
  }
  __syncthreads();
  const double hk = hs;
  double s = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const T ri = (T)((double)r[i] - hk * (double)v[i]);
    r[i] = ri;
    s += (double)ri * (double)(w ? w[i] : ri);
  }
  __syncthreads();
  s = block_sum(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T>
__global__ void k_scale(int n, double a, const double* __restrict__ a_dev, int a_dev_inv,
                        const T* __restrict__ x, T* __restrict__ y) {
  double aa = a;
  if (a_dev) aa = a_dev_inv ? 1.0 / a_dev[0] : a_dev[0];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = (T)(aa * (double)x[i]);
}

template <typename TO, typename TI>
__global__ void k_convert(int n, const TI* __restrict__ x, TO* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = (TO)x[i];
}

__global__ void k_mul(int n, const double* __restrict__ a, const double* __restrict__ x,
                      double* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = a[i] * x[i];
}

template <typename T>
__global__ void k_fill(int n, T v, T* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = v;
}

// y[i] += a*x[i] for i in [from, n)
template <typename TY, typename TX>
__global__ void k_add_scaled_range(int from, int n, double a, const TX* __restrict__ x,
                                   TY* __restrict__ y) {
  for (int i = from + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    y[i] = (TY)((double)y[i] + a * (double)x[i]);
}

// dinv[i] = 1 / A_ii ; flags zero diagonals.  16 lanes per row read the
// row's entries coalesced; the one lane holding the diagonal entry (a sum
// with zeros from the others, exact) gives d
__global__ __launch_bounds__(256) void k_inv_diag(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const double* __restrict__ va, double* __restrict__ dinv,
                                                  int* __restrict__ zero_flag) {
  const int l = threadIdx.x & 15;
  for (int r = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 4); r < n; r += (int)((gridDim.x * blockDim.x) >> 4)) {
    double d = 0;
    for (int k = rp[r] + l; k < rp[r + 1]; k += 16)
      if (ci[k] == r) d = va[k];
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) d += __shfl_xor(d, m, 16);
    if (l == 0) {
      if (d == 0.0) { atomicOr(zero_flag, 1); dinv[r] = 0.0; }
      else dinv[r] = 1.0 / d;
    }
  }
}

// FP64 values -> bf16 values with an FP32 per-row scale (row max-abs
// normalised to 1), for the mixed-precision V-cycle
// per-row scale max |a_ij| (16 lanes per row, coalesced; max is exact in
// any order) and, when vb is given, the row's bf16 values a_ij / scale
__global__ __launch_bounds__(256) void k_to_bf16_rows(int n, const int* __restrict__ rp, const double* __restrict__ va,
                                                      uint16_t* __restrict__ vb, float* __restrict__ rscale) {
  const int l = threadIdx.x & 15;
  for (int r = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 4); r < n; r += (int)((gridDim.x * blockDim.x) >> 4)) {
    const int k0 = rp[r], k1 = rp[r + 1];
    double m = 0;
    for (int k = k0 + l; k < k1; k += 16) m = fmax(m, fabs(va[k]));
#pragma unroll
    for (int s = 8; s >= 1; s >>= 1) m = fmax(m, __shfl_xor(m, s, 16));
    const double sc = (m > 0) ? m : 1.0;
    if (l == 0) rscale[r] = (float)sc;
    if (!vb) continue;   // row scales only
    const double inv = 1.0 / sc;
    for (int k = k0 + l; k < k1; k += 16) {
      const float f = (float)(va[k] * inv);
      uint32_t u = __float_as_uint(f);
      u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
      vb[k] = (uint16_t)(u >> 16);
    }
  }
}

int grid_for(int n, int block = 256, int cap = 4096) {
  int g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return g > cap ? cap : g;
}

}  // namespace

// ----------------------------------------------------------------- launchers
void fe_apply_scaled_add(const DevCSR& A, const double* x, double* y, double a, hipStream_t st) {
  dispatch_fe_spmv<MODE_SCALED_ADD, double, double>(A, A.val, nullptr, x, nullptr, nullptr, y, a,
                                                    nullptr, st);
}
void fe_apply_scaled_add3(const DevCSR& A, const double* x, double* y, size_t ld, double a, hipStream_t st) {
  if (A.sell.n_slices == 0 || !A.sell.val || A.sell.off || A.lpr != 0) {
    for (int k = 0; k < 3; ++k) fe_apply_scaled_add(A, x + k * ld, y + k * ld, a, st);
    return;
  }
  if (A.n < A.sell3_min_rows) {
    // too few rows for one lane per row (a coarse multigrid level: ~80
    // entries per row summed one after another take as long as the finest
    // level's launch); 8 lanes per row of the CSR values instead
    for (int k = 0; k < 3; ++k)
      launch_fe_spmv<8, MODE_SCALED_ADD>(A, A.val, nullptr, x + k * ld, (const double*)nullptr,
                                         (const double*)nullptr, y + k * ld, a, nullptr, st);
    return;
  }
  int sb = 0;
  const int grid = sell_grid(A, &sb);
  hipLaunchKernelGGL(k_sell_spmv3, dim3(grid), dim3(256), 0, st, A.sell.n_slices, sb, A.n, A.active, ld,
                     A.sell.slice_ptr, A.sell.row, A.sell.col, A.sell.val, x, y, a);
}
void fe_multiply_active(const DevCSR& A, const double* x, double* y, double a, hipStream_t st) {
  DevCSR Aa = A;
  Aa.n = A.active;   // no identity rows
  dispatch_fe_spmv<MODE_SCALED_ADD, double, double>(Aa, A.val, nullptr, x, nullptr, nullptr, y, a,
                                                    nullptr, st);
}
void fe_apply(const DevCSR& A, const double* x, double* y, hipStream_t st) {
  dispatch_fe_spmv<MODE_APPLY, double, double>(A, A.val, nullptr, x, nullptr, nullptr, y, 1.0,
                                               nullptr, st);
}
// d = b - A x with FP32 copies of the values, FP64 vectors and sums (CSR
// kernel with A.lpr lanes per row; no partial sums kept)
void fe_defect_f32val(const DevCSR& A, const float* val32, const double* x, const double* b, double* d,
                      double* partial, hipStream_t st) {
  switch (A.lpr) {
    case 4: launch_fe_spmv<4, MODE_DEFECT>(A, val32, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
    case 16: launch_fe_spmv<16, MODE_DEFECT>(A, val32, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
    case 32: launch_fe_spmv<32, MODE_DEFECT>(A, val32, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
    default: launch_fe_spmv<8, MODE_DEFECT>(A, val32, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
  }
}
// the same with bf16 copies of the values
void fe_defect_bf16val(const DevCSR& A, const uint16_t* val16, const double* x, const double* b, double* d,
                       double* partial, hipStream_t st) {
  switch (A.lpr) {
    case 4: launch_fe_spmv<4, MODE_DEFECT>(A, val16, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
    case 16: launch_fe_spmv<16, MODE_DEFECT>(A, val16, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
    case 32: launch_fe_spmv<32, MODE_DEFECT>(A, val16, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
    default: launch_fe_spmv<8, MODE_DEFECT>(A, val16, nullptr, x, b, (const double*)nullptr, d, 1.0, partial, st); break;
  }
}
// number of blocks (= partial sums) the FE-matrix kernel launches for A
static int spmv_grid(const DevCSR& A, bool bf16) {
  const bool have_sell = bf16 ? (A.sell.val_bf16 != nullptr) : (A.sell.val != nullptr);
  if (have_sell && A.lpr == 0) {
    int sb = 0;
    return sell_grid(A, &sb);
  }
  const int lpr = A.lpr ? A.lpr : 64;
  return (A.n + (256 / lpr) - 1) / (256 / lpr);
}

void fe_defect(const DevCSR& A, const double* x, const double* b, double* d, double* partial,
               double* norm_out, hipStream_t st) {
  dispatch_fe_spmv<MODE_DEFECT, double, double>(A, A.val, nullptr, x, b, nullptr, d, 1.0, partial,
                                                st);
  const int nb = spmv_grid(A, false);
  if (norm_out) hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, nb, partial, norm_out, 1);
}
// upper bound of the partial sums any FE-matrix kernel writes for A
int fe_defect_partials(const DevCSR& A) {
  const int lpr = A.lpr ? A.lpr : 64;
  const int csr = (A.n + (256 / lpr) - 1) / (256 / lpr);
  const int csr64 = (A.n + 3) / 4;
  int sb = 0;
  const int sell = A.sell.n_slices ? sell_grid(A, &sb) : 0;
  int m = csr > sell ? csr : sell;
  return m > csr64 ? m : csr64;
}
void fe_jacobi(const DevCSR& A, const double* x, const double* b, const double* dinv, double* xn,
               double omega, hipStream_t st) {
  dispatch_fe_spmv<MODE_JACOBI, double, double>(A, A.val, nullptr, x, b, dinv, xn, omega, nullptr,
                                                st);
}

template <typename VT, typename XT>
static bool jacobi_from_zero(const DevCSR& A, const VT* bd_val, const float* rscale, const XT* x,
                             const XT* b, const XT* dinv, XT* xn, double omega, hipStream_t st) {
  if (A.lpr != 0 || !A.sell.bd_ptr || (!bd_val && A.sell.n_bd > 0)) return false;
  if (A.n == 0) return true;
  if (A.sell.dir_zero) {
    // x vanishes everywhere: x' = w D^-1 b (the row sums are exactly zero)
    hipLaunchKernelGGL((k_jacobi_from_zero_nodir<XT>), dim3((A.n + 255) / 256), dim3(256), 0, st, A.n, b, dinv, xn,
                       omega);
    return true;
  }
  XT* g = nullptr;
  if (A.sell.off && A.sell.gsum) {
    if (A.sell.gsum_bytes != (int)sizeof(XT)) throw std::runtime_error("folded SELL layout: sums of another precision");
    g = (XT*)A.sell.gsum;
  }
  hipLaunchKernelGGL((k_sell_jacobi_from_zero<VT, XT>), dim3((A.n + 255) / 256), dim3(256), 0, st, A.n,
                     A.active, A.sell.bd_ptr, A.sell.bd_col, bd_val, rscale, x, b, dinv, xn, omega, g);
  return true;
}

template <typename VT, typename XT>
static void dirichlet_sums(const DevCSR& A, const VT* bd_val, const XT* x, hipStream_t st) {
  if (!(A.sell.off && A.sell.gsum) || A.active <= 0 || A.sell.dir_zero) return;
  if (A.sell.gsum_bytes != (int)sizeof(XT)) throw std::runtime_error("folded SELL layout: sums of another precision");
  if (A.sell.n_bd > 0 && (!bd_val || !A.sell.bd_ptr)) throw std::runtime_error("folded SELL layout without Dirichlet values");
  hipLaunchKernelGGL((k_sell_dirichlet_sums<VT, XT>), dim3((A.active + 255) / 256), dim3(256), 0, st, A.active,
                     A.sell.bd_ptr, A.sell.bd_col, bd_val, x, (XT*)A.sell.gsum);
}
void sell_dirichlet_sums(const DevCSR& A, const double* x, hipStream_t st) {
  dirichlet_sums<double, double>(A, A.sell.bd_val, x, st);
}
void sell_dirichlet_sums_bf16(const DevCSR& A, const float* x, hipStream_t st) {
  dirichlet_sums<uint16_t, float>(A, A.sell.bd_val_bf16, x, st);
}
bool fe_jacobi_from_zero(const DevCSR& A, const double* x, const double* b, const double* dinv, double* xn,
                         double omega, hipStream_t st) {
  return jacobi_from_zero<double, double>(A, A.sell.bd_val, nullptr, x, b, dinv, xn, omega, st);
}
bool fe_jacobi_from_zero_bf16(const DevCSR& A, const float* x, const float* b, const float* dinv, float* xn,
                              double omega, hipStream_t st) {
  return jacobi_from_zero<uint16_t, float>(A, A.sell.bd_val_bf16, A.rscale, x, b, dinv, xn, omega, st);
}
// one thread per active row: its Dirichlet-column values (FP64 copy and/or
// bf16 of value / row scale, rounded to nearest even as in k_sell_fill_bf16)
__global__ void k_sell_dirichlet_fill(int na, const int* __restrict__ bd_ptr, const int* __restrict__ bd_src,
                                      const double* __restrict__ v, const float* __restrict__ rscale,
                                      double* __restrict__ out64, uint16_t* __restrict__ out16) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= na) return;
  const int e = bd_ptr[r + 1];
  const double inv = rscale ? 1.0 / (double)rscale[r] : 0.0;
  for (int k = bd_ptr[r]; k < e; ++k) {
    const double a = v[bd_src[k]];
    if (out64) out64[k] = a;
    if (out16) {
      const float f = (float)(a * inv);
      uint32_t u = __float_as_uint(f);
      u += 0x7FFFu + ((u >> 16) & 1u);
      out16[k] = (uint16_t)(u >> 16);
    }
  }
}
void sell_dirichlet_fill(const DevCSR& A, double* out64, uint16_t* out16, hipStream_t st) {
  if (A.sell.n_bd <= 0 || !A.sell.bd_ptr || A.active <= 0) return;
  hipLaunchKernelGGL(k_sell_dirichlet_fill, dim3((A.active + 255) / 256), dim3(256), 0, st, A.active,
                     A.sell.bd_ptr, A.sell.bd_src, A.val, out16 ? A.rscale : nullptr, out64, out16);
}

template <typename S, typename D>
__global__ void k_gather(int n, const int* __restrict__ perm, const S* __restrict__ x, D* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (D)x[perm[i]];
}
template <typename S, typename D>
__global__ void k_scatter(int n, const int* __restrict__ perm, const S* __restrict__ x, D* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[perm[i]] = (D)x[i];
}
template <typename S, typename D>
static void gather_t(int n, const int* perm, const S* x, D* y, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL((k_gather<S, D>), dim3((n + 255) / 256), dim3(256), 0, st, n, perm, x, y);
}
template <typename S, typename D>
static void scatter_t(int n, const int* perm, const S* x, D* y, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL((k_scatter<S, D>), dim3((n + 255) / 256), dim3(256), 0, st, n, perm, x, y);
}
void vec_gather(int n, const int* perm, const double* x, double* y, hipStream_t st) { gather_t(n, perm, x, y, st); }
void vec_gather(int n, const int* perm, const double* x, float* y, hipStream_t st) { gather_t(n, perm, x, y, st); }
void vec_gather(int n, const int* perm, const float* x, float* y, hipStream_t st) { gather_t(n, perm, x, y, st); }
void vec_scatter(int n, const int* perm, const double* x, double* y, hipStream_t st) { scatter_t(n, perm, x, y, st); }
void vec_scatter(int n, const int* perm, const float* x, double* y, hipStream_t st) { scatter_t(n, perm, x, y, st); }

// mixed precision: bf16 values with FP32 row scale, FP32 vectors
void fe_apply_scaled_add_bf16(const DevCSR& A, const float* x, float* y, double a, hipStream_t st) {
  dispatch_fe_spmv<MODE_SCALED_ADD, uint16_t, float>(A, A.val_bf16, A.rscale, x, nullptr, nullptr,
                                                     y, a, nullptr, st);
}
void fe_defect_bf16(const DevCSR& A, const float* x, const float* b, float* d, double* partial,
                    double* norm_out, hipStream_t st) {
  dispatch_fe_spmv<MODE_DEFECT, uint16_t, float>(A, A.val_bf16, A.rscale, x, b, nullptr, d, 1.0,
                                                 partial, st);
  const int nb = spmv_grid(A, true);
  if (norm_out) hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, nb, partial, norm_out, 1);
}
void fe_jacobi_bf16(const DevCSR& A, const float* x, const float* b, const float* dinv, float* xn,
                    double omega, hipStream_t st) {
  dispatch_fe_spmv<MODE_JACOBI, uint16_t, float>(A, A.val_bf16, A.rscale, x, b, dinv, xn, omega,
                                                 nullptr, st);
}

template <typename T>
void csr_spmv_t(const DevCSR& P, const T* x, T* y, double alpha, double beta, int row_end,
                int zero_rest, hipStream_t st) {
  if (P.pt.n_slices > 0 && P.sell.n_rows == row_end) {
    // pattern slices, the SELL copy of the other rows and the zero tail in one launch
    const int pb = ((P.pt.n_slices + 3) / 4 + 7) / 8 * 8;
    const int sb = ((P.sell.n_slices + 3) / 4 + 7) / 8 * 8;
    const int tail = zero_rest ? (P.n - row_end + 255) / 256 : 0;
    auto go = [&](auto kern, const auto* va) {
      hipLaunchKernelGGL(kern, dim3(pb + sb + tail), dim3(256), 0, st, P.pt.n_slices, pb, P.pt.row, P.pt.base,
                         P.pt.slice_pat, P.pt.pat_ptr, P.pt.pat_col, P.pt.pat_val, P.sell.n_slices, sb, P.n, row_end,
                         P.sell.slice_ptr, P.sell.row, P.sell.col, va, x, y, alpha, beta, zero_rest);
    };
    if (P.sell.val_bf16) go(k_pat_transfer<T, uint16_t>, (const uint16_t*)P.sell.val_bf16);
    else go(k_pat_transfer<T, double>, (const double*)P.sell.val);
    return;
  }
  if (P.sell.n_slices > 0 && P.sell.n_rows == row_end) {
    const int sb = ((P.sell.n_slices + 3) / 4 + 7) / 8 * 8;
    const int tail = zero_rest ? (P.n - row_end + 255) / 256 : 0;
    const bool h = P.sell.val_bf16 != nullptr;
    if (h)
      hipLaunchKernelGGL((k_sell_transfer<T, uint16_t>), dim3(sb + tail), dim3(256), 0, st, P.sell.n_slices, sb, P.n,
                         row_end, P.sell.slice_ptr, P.sell.row, P.sell.col, P.sell.val_bf16, x, y, alpha, beta,
                         zero_rest);
    else
      hipLaunchKernelGGL((k_sell_transfer<T, double>), dim3(sb + tail), dim3(256), 0, st, P.sell.n_slices, sb, P.n,
                         row_end, P.sell.slice_ptr, P.sell.row, P.sell.col, P.sell.val, x, y, alpha, beta,
                         zero_rest);
    return;
  }
  // lanes per row: P.lpr (1, 2, 4 or 8; short interpolation rows use fewer)
  const int lpr = (P.lpr == 1 || P.lpr == 2 || P.lpr == 4) ? P.lpr : 8;
  const int grid = (P.n + (256 / lpr) - 1) / (256 / lpr);
  if (grid == 0) return;
  auto go = [&](auto kern, const auto* va) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, P.n, P.row_ptr, P.col, va, x, y, alpha, beta, row_end,
                       zero_rest);
  };
  // exact bf16 copy of the values (val_bf16 without row scales), else FP64
  const bool h = P.val_bf16 && !P.rscale;
  switch (lpr) {
    case 1: h ? go(k_csr_spmv<1, T, uint16_t>, P.val_bf16) : go(k_csr_spmv<1, T, double>, P.val); break;
    case 2: h ? go(k_csr_spmv<2, T, uint16_t>, P.val_bf16) : go(k_csr_spmv<2, T, double>, P.val); break;
    case 4: h ? go(k_csr_spmv<4, T, uint16_t>, P.val_bf16) : go(k_csr_spmv<4, T, double>, P.val); break;
    default: h ? go(k_csr_spmv<8, T, uint16_t>, P.val_bf16) : go(k_csr_spmv<8, T, double>, P.val);
  }
}
void csr_b3(const DevCSR& B, const double* v2, const double* v3, const double* x, size_t ldx, double* y, double a,
            int row_end, hipStream_t st) {
  const int grid = (B.n + 15) / 16;   // 16 lanes per row (rows of ~125 entries)
  if (grid > 0)
    hipLaunchKernelGGL(k_csr_b3<16>, dim3(grid), dim3(256), 0, st, B.n, B.row_ptr, B.col, B.val, v2, v3, x, ldx, y,
                       a, row_end);
}
void csr_bt3(const DevCSR& T, const double* v2, const double* v3, const double* p, double* y, size_t ldy, double a,
             int row_end, hipStream_t st) {
  const int grid = (T.n + 31) / 32;
  if (grid > 0)
    hipLaunchKernelGGL(k_csr_bt3<8>, dim3(grid), dim3(256), 0, st, T.n, T.row_ptr, T.col, T.val, v2, v3, p, y, ldy,
                       a, row_end);
}
void csr_spmv3(const DevCSR& P, const double* x, size_t ldx, double* y, size_t ldy, double alpha, double beta,
               int row_end, int zero_rest, hipStream_t st) {
  if (P.sell.n_slices > 0) {   // the SELL transfer copy: one vector at a time
    for (int k = 0; k < 3; ++k) csr_spmv(P, x + k * ldx, y + k * ldy, alpha, beta, row_end, zero_rest, st);
    return;
  }
  const int lpr = (P.lpr == 1 || P.lpr == 2 || P.lpr == 4) ? P.lpr : 8;
  const int grid = (P.n + (256 / lpr) - 1) / (256 / lpr);
  if (grid == 0) return;
  const bool h = P.val_bf16 && !P.rscale;
  auto go = [&](auto kern, const auto* va) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, P.n, P.row_ptr, P.col, va, x, ldx, y, ldy, alpha, beta,
                       row_end, zero_rest);
  };
  switch (lpr) {
    case 1: h ? go(k_csr_spmv3<1, uint16_t>, P.val_bf16) : go(k_csr_spmv3<1, double>, P.val); break;
    case 2: h ? go(k_csr_spmv3<2, uint16_t>, P.val_bf16) : go(k_csr_spmv3<2, double>, P.val); break;
    case 4: h ? go(k_csr_spmv3<4, uint16_t>, P.val_bf16) : go(k_csr_spmv3<4, double>, P.val); break;
    default: h ? go(k_csr_spmv3<8, uint16_t>, P.val_bf16) : go(k_csr_spmv3<8, double>, P.val);
  }
}
void csr_spmv(const DevCSR& P, const double* x, double* y, double alpha, double beta, int row_end,
              int zero_rest, hipStream_t st) {
  csr_spmv_t<double>(P, x, y, alpha, beta, row_end, zero_rest, st);
}
void csr_spmv_f32(const DevCSR& P, const float* x, float* y, double alpha, double beta,
                  int row_end, int zero_rest, hipStream_t st) {
  csr_spmv_t<float>(P, x, y, alpha, beta, row_end, zero_rest, st);
}

void vec_dot(int n, const double* x, const double* y, double* partial, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial<double>, dim3(kRedBlocks), dim3(256), 0, st, n, x, y, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, kRedBlocks, partial, out, 0);
}
void vec_norm(int n, const double* x, double* partial, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial<double>, dim3(kRedBlocks), dim3(256), 0, st, n, x, x, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, kRedBlocks, partial, out, 1);
}
void vec_norm_f32(int n, const float* x, double* partial, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial<float>, dim3(kRedBlocks), dim3(256), 0, st, n, x, x, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, kRedBlocks, partial, out, 1);
}
void vec_axpy_dot(int n, const double* h, const double* v, double* r, const double* w,
                  double* partial, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dot, dim3(kRedBlocks), dim3(256), 0, st, n, h, v, r, w, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, kRedBlocks, partial, out,
                     w ? 0 : 1);
}
int vec_partials() { return kRedBlocks; }
void vec_dot_partial(int n, const double* x, const double* y, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial<double>, dim3(kRedBlocks), dim3(256), 0, st, n, x, y, partial);
}
void vec_axpy_dot_reduced(int n, const double* partial_in, double* h_out, const double* v, double* r,
                          const double* w, double* partial_out, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dot_reduced<double>, dim3(kRedBlocks), dim3(256), 0, st, n, kRedBlocks, partial_in, h_out, v,
                     r, w, partial_out);
}
void vec_reduce(const double* partial, double* out, int sqrt_of_sum, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, kRedBlocks, partial, out, sqrt_of_sum);
}
void vec_axpy(int n, double a, const double* a_dev, double sign, const double* x, double* y,
              hipStream_t st) {
  hipLaunchKernelGGL(k_axpy<double>, dim3(grid_for(n)), dim3(256), 0, st, n, a, a_dev, sign, x, y);
}
void vec_axpy_f32(int n, double a, const float* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy<float>, dim3(grid_for(n)), dim3(256), 0, st, n, a, (const double*)nullptr,
                     1.0, x, y);
}
void vec_scale(int n, double a, const double* a_dev, int a_dev_inv, const double* x, double* y,
               hipStream_t st) {
  hipLaunchKernelGGL(k_scale<double>, dim3(grid_for(n)), dim3(256), 0, st, n, a, a_dev, a_dev_inv, x, y);
}
void vec_scale_f32(int n, double a, const double* a_dev, int a_dev_inv, const float* x, float* y,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_scale<float>, dim3(grid_for(n)), dim3(256), 0, st, n, a, a_dev, a_dev_inv, x, y);
}
void vec_mul(int n, const double* a, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(256), 0, st, n, a, x, y);
}
void vec_fill(int n, double v, double* y, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(256), 0, st, n, v, y);
}
void vec_fill_f32(int n, float v, float* y, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(256), 0, st, n, v, y);
}
void vec_add_scaled_range(int from, int n, double a, const double* x, double* y, hipStream_t st) {
  if (n > from)
    hipLaunchKernelGGL((k_add_scaled_range<double, double>), dim3(grid_for(n - from)), dim3(256), 0, st,
                       from, n, a, x, y);
}
void vec_add_scaled_range_f32(int from, int n, double a, const float* x, float* y, hipStream_t st) {
  if (n > from)
    hipLaunchKernelGGL((k_add_scaled_range<float, float>), dim3(grid_for(n - from)), dim3(256), 0, st,
                       from, n, a, x, y);
}
void vec_f64_to_f32(int n, const double* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
}
void vec_f32_to_f64(int n, const float* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
}
void csr_inv_diag(const DevCSR& A, double* dinv, int* zero_flag, hipStream_t st) {
  hipLaunchKernelGGL(k_inv_diag, dim3(grid_for(A.n, 16)), dim3(256), 0, st, A.n, A.row_ptr, A.col, A.val,
                     dinv, zero_flag);
}
void sell_fill_f64(const DevCSR& A, double* sell_val, hipStream_t st) {
  if (A.sell.total)
    hipLaunchKernelGGL(k_sell_fill<double>, dim3(4096), dim3(256), 0, st, A.sell.total, A.sell.src, A.val, sell_val);
}
void sell_fill_bf16(const DevCSR& A, uint16_t* sell_bf16, hipStream_t st) {
  if (A.sell.n_slices)
    hipLaunchKernelGGL(k_sell_fill_bf16, dim3((A.sell.n_slices + 3) / 4), dim3(256), 0, st, A.sell.n_slices,
                       A.sell.slice_ptr, A.sell.row, A.sell.src, A.val, A.rscale, sell_bf16, A.sell.pack);
}
void csr_to_bf16(const DevCSR& A, uint16_t* vb, float* rscale, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16_rows, dim3(grid_for(A.n, 16)), dim3(256), 0, st, A.n, A.row_ptr, A.val, vb,
                     rscale);
}

// ------------------------------------------- FP32 Krylov of the refinement
namespace {

// y = x / a_dev[0], rounded to FP32
__global__ void k_scale_to_f32(int n, const double* __restrict__ x, const double* __restrict__ a_dev,
                               float* __restrict__ y) {
  const double s = 1.0 / a_dev[0];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = (float)(s * x[i]);
}

struct Coef64 {
  double c[64];
};
// x += sum_j c_j Z_j (Z_j = Z + j ld, FP32), summed in j order in FP64 per entry
__global__ void k_multi_axpy_f32(int n, int k, Coef64 c, const float* __restrict__ Z, size_t ld,
                                 double* __restrict__ x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0;
    for (int j = 0; j < k; ++j) s += c.c[j] * (double)Z[(size_t)j * ld + i];
    x[i] += s;
  }
}

}  // namespace

void sell_fill_f32(const DevCSR& A, float* out, hipStream_t st) {
  if (A.sell.total)
    hipLaunchKernelGGL(k_sell_fill<float>, dim3(4096), dim3(256), 0, st, A.sell.total, A.sell.src, A.val, out);
}
void vec_dot_partial_f32(int n, const float* x, const float* y, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial<float>, dim3(kRedBlocks), dim3(256), 0, st, n, x, y, partial);
}
void vec_axpy_dot_reduced_f32(int n, const double* partial_in, double* h_out, const float* v, float* r,
                              const float* w, double* partial_out, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dot_reduced<float>, dim3(kRedBlocks), dim3(256), 0, st, n, kRedBlocks, partial_in, h_out,
                     v, r, w, partial_out);
}
// ---------------- classical Gram-Schmidt over an FP32 basis (FP64 sums)
namespace {
constexpr int kMaxDots = 32;

// partial[b kMaxDots + k] = sum over block b's entries of w_i V_k,i (k < NV).
// Four consecutive entries per thread and step (16-byte loads; ld and the
// vectors 16-byte aligned), the last n % 4 entries by block 0.
typedef float v4f_cgs __attribute__((ext_vector_type(4)));
template <int NV>
__global__ __launch_bounds__(256) void k_multi_dot_f32(int n, const float* __restrict__ w,
                                                       const float* __restrict__ V, size_t ld,
                                                       double* __restrict__ partial) {
  __shared__ double red[4][kMaxDots];
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  const int n4 = n >> 2;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x) {
    const v4f_cgs wi = ((const v4f_cgs*)w)[i];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const v4f_cgs v = ((const v4f_cgs*)(V + (size_t)k * ld))[i];
      acc[k] += ((double)wi.x * (double)v.x + (double)wi.y * (double)v.y) +
                ((double)wi.z * (double)v.z + (double)wi.w * (double)v.w);
    }
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < (n & 3)) {
    const int i = 4 * n4 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += (double)w[i] * (double)V[(size_t)k * ld + i];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double t = group_sum<64>(acc[k]);
    if (lane == 0) red[wv][k] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x < NV)
    partial[(size_t)blockIdx.x * kMaxDots + threadIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// out[k] = sum over the m blocks of partial[b kMaxDots + k], in a fixed order
__global__ __launch_bounds__(1024) void k_multi_reduce(int m, int nv, const double* __restrict__ partial,
                                                       double* __restrict__ out) {
  __shared__ double red[32][kMaxDots + 1];
  const int k = threadIdx.x & 31, g = threadIdx.x >> 5;
  double s = 0.0;
  if (k < nv)
    for (int b = g; b < m; b += 32) s += partial[(size_t)b * kMaxDots + k];
  red[g][k] = s;
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    double t = 0.0;
    for (int q = 0; q < 32; ++q) t += red[q][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// w -= sum_k h_k V_k (in FP64, rounded once), partial[b] = block b's part of
// |w|^2; four entries per thread and step as in k_multi_dot_f32
template <int NV>
__global__ __launch_bounds__(256) void k_multi_axpy_norm_f32(int n, const double* __restrict__ h,
                                                             const float* __restrict__ V, size_t ld,
                                                             float* __restrict__ w, double* __restrict__ partial) {
  __shared__ double hs[kMaxDots];
  __shared__ double red[16];
  if ((int)threadIdx.x < NV) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  double s2 = 0.0;
  const int n4 = n >> 2;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x) {
    const v4f_cgs wi = ((const v4f_cgs*)w)[i];
    double a0 = wi.x, a1 = wi.y, a2 = wi.z, a3 = wi.w;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const v4f_cgs v = ((const v4f_cgs*)(V + (size_t)k * ld))[i];
      const double hk = hs[k];
      a0 -= hk * (double)v.x;
      a1 -= hk * (double)v.y;
      a2 -= hk * (double)v.z;
      a3 -= hk * (double)v.w;
    }
    v4f_cgs o;
    o.x = (float)a0; o.y = (float)a1; o.z = (float)a2; o.w = (float)a3;
    ((v4f_cgs*)w)[i] = o;
    s2 += ((double)o.x * (double)o.x + (double)o.y * (double)o.y) + ((double)o.z * (double)o.z + (double)o.w * (double)o.w);
  }
  if (blockIdx.x == 0 && (int)threadIdx.x < (n & 3)) {
    const int i = 4 * n4 + threadIdx.x;
    double a = (double)w[i];
#pragma unroll
    for (int k = 0; k < NV; ++k) a -= hs[k] * (double)V[(size_t)k * ld + i];
    const float wi = (float)a;
    w[i] = wi;
    s2 += (double)wi * (double)wi;
  }
  s2 = block_sum(s2, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s2;
}
}  // namespace

namespace {
// host_dst[0..cnt) = src, then *host_flag = seq with system-scope release, so
// a host thread that sees the flag (acquire) sees the values
__global__ __launch_bounds__(64) void k_publish_to_host(const double* __restrict__ src, int cnt,
                                                        double* host_dst, unsigned int* host_flag, unsigned int seq) {
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) host_dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(host_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
}  // namespace

void publish_to_host(const double* src, int cnt, double* host_dst, unsigned int* host_flag, unsigned int seq,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_publish_to_host, dim3(1), dim3(64), 0, st, src, cnt, host_dst, host_flag, seq);
}

int vec_multi_dot_max() { return kMaxDots; }
int vec_multi_partials() { return kRedBlocks * kMaxDots; }
namespace {
template <int NV>
void cgs_nv(int n, const float* V, size_t ld, float* w, double* h_out, double* mpartial, double* partial,
            hipStream_t st) {
  hipLaunchKernelGGL(k_multi_dot_f32<NV>, dim3(kRedBlocks), dim3(256), 0, st, n, w, V, ld, mpartial);
  hipLaunchKernelGGL(k_multi_reduce, dim3(1), dim3(1024), 0, st, kRedBlocks, NV, mpartial, h_out);
  hipLaunchKernelGGL(k_multi_axpy_norm_f32<NV>, dim3(kRedBlocks), dim3(256), 0, st, n, h_out, V, ld, w, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, st, kRedBlocks, partial, h_out + NV, 1);
}
template <int NV>
void cgs_dispatch(int nv, int n, const float* V, size_t ld, float* w, double* h_out, double* mpartial,
                  double* partial, hipStream_t st) {
  if constexpr (NV <= kMaxDots) {
    if (nv == NV) cgs_nv<NV>(n, V, ld, w, h_out, mpartial, partial, st);
    else cgs_dispatch<NV + 1>(nv, n, V, ld, w, h_out, mpartial, partial, st);
  }
}
}  // namespace

void vec_cgs_f32(int n, int nv, const float* V, size_t ld, float* w, double* h_out, double* mpartial,
                 double* partial, hipStream_t st) {
  if (nv < 1 || nv > kMaxDots) throw std::runtime_error("classical Gram-Schmidt: too many basis vectors");
  if (ld % 4 != 0 || ((uintptr_t)V | (uintptr_t)w) % 16 != 0)
    throw std::runtime_error("classical Gram-Schmidt: basis vectors must be 16-byte aligned");
  cgs_dispatch<1>(nv, n, V, ld, w, h_out, mpartial, partial, st);
}

void vec_scale_to_f32(int n, const double* x, const double* a_dev, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_to_f32, dim3(grid_for(n)), dim3(256), 0, st, n, x, a_dev, y);
}
void vec_multi_axpy_f32(int n, int k, const double* c, const float* Z, size_t ld, double* x, hipStream_t st) {
  if (k < 1 || k > 64) throw std::runtime_error("vec_multi_axpy_f32: 1 <= k <= 64");
  Coef64 cc{};
  for (int j = 0; j < k; ++j) cc.c[j] = c[j];
  hipLaunchKernelGGL(k_multi_axpy_f32, dim3(grid_for(n)), dim3(256), 0, st, n, k, cc, Z, ld, x);
}
template <typename VT>
static void apply_active(const DevCSR& A, const VT* vals, const float* rscale, const float* x, float* y,
                         const float* zero_gsum, hipStream_t st) {
  if (!A.sell.off || (A.sell.pack != 1 && !A.sell.slice_pat) || A.sell.n_slices == 0)
    throw std::runtime_error("the refinement operator needs the folded 16-bit SELL layout");
  int sb = 0;
  const int grid = sell_grid(A, &sb);
#define PMX_PATA(PK, C)                                                                                              \
  hipLaunchKernelGGL((k_sell_pat<MODE_APPLY, VT, float, false, PK, 4, C>), dim3(grid), dim3(256), 0, st,           \
                     A.sell.n_slices, sb, A.n, A.active, A.sell.slice_ptr, A.sell.row, A.sell.slice_pat,            \
                     A.sell.pat_off, vals, rscale, x, (const float*)nullptr, (const float*)nullptr, y, 1.0,          \
                     (double*)nullptr, (const float*)nullptr)
  if (A.sell.slice_pat && A.sell.gsum) {
    if (A.sell.pack == 4) PMX_PATA(4, 16);
    else PMX_PATA(1, 16);
  }
#undef PMX_PATA
  else
    hipLaunchKernelGGL((k_sell_spmv<MODE_APPLY, VT, float, true, true>), dim3(grid), dim3(256), 0, st,
                       A.sell.n_slices, sb, A.n, A.active, A.sell.slice_ptr, A.sell.row, A.sell.col, A.sell.off,
                       A.sell.bd_ptr, A.sell.bd_col, (const VT*)nullptr, vals, rscale, x, (const float*)nullptr,
                       (const float*)nullptr, y, 1.0, (double*)nullptr, zero_gsum);
}
void fe_apply_active_f32(const DevCSR& A, const float* vals, const float* x, float* y, const float* zero_gsum,
                         hipStream_t st) {
  apply_active<float>(A, vals, nullptr, x, y, zero_gsum, st);
}
void fe_apply_active_bf16(const DevCSR& A, const float* x, float* y, const float* zero_gsum, hipStream_t st) {
  if (!A.sell.val_bf16 || !A.rscale) throw std::runtime_error("no bf16 SELL values");
  apply_active<uint16_t>(A, A.sell.val_bf16, A.rscale, x, y, zero_gsum, st);
}

}  // namespace pmx

// ------------------------------------------------------------------
// Dense in-place inversion by blocked Gauss-Jordan elimination with partial
// pivoting (set-up of the exact coarse-grid solve); A is n x n row-major.
// Per block column K of width b <= 64:
//  0. speculation: k_bgj_block_inverse (spec) and k_bgj_check_below test
//     whether partial pivoting on the panel A[K:, K:K+b] would keep the
//     diagonal pivots (every multiplier at most 1 in magnitude, ties to the
//     lower row); then steps 1, 2 and the second block inverse do nothing;
//  1. k_bgj_panel_reg (one workgroup; k_bgj_panel when the panel has more
//     than 5,120 rows): LU with partial pivoting of the panel
//     in a scratch copy chooses the b pivot rows;
//  2. k_bgj_swap: the pivot rows are swapped into rows K..K+b-1;
//  3. the block sweep with P = A_KK^-1 (k_bgj_block_inverse, Gauss-Jordan
//     without further pivoting on the b x b block, whose LU needs none after
//     step 1), in place (o: the rows / columns outside the block):
//        C = A[o, K] (compact copy, k_bgj_make_c)
//        R = A[K, o] = P A[K, o],  A[K, K] = P      (k_bgj_rows)
//        A[o, o] -= C R                              (k_bgj_update)
//        A[o, K] = -A[o, K] P                        (k_bgj_cols)
//     the products on FP64 matrix cores.
// The row swaps leave the inverse of the row-permuted matrix; its columns
// are put back in order at the end (k_bgj_unpermute, in place, one row per
// workgroup through LDS).  The trailing update
// is a rank-b GEMM per block instead of b rank-1 sweeps over the matrix.
namespace pmx {
namespace {

constexpr int kBgjBlock = 64;
typedef double v4d_bgj __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(1024) void k_bgj_panel(int n, int K, int b, const double* __restrict__ A,
                                                     double* __restrict__ W, int* __restrict__ piv,
                                                     int* __restrict__ singular, const int* __restrict__ need) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  if (!*need) return;   // the diagonal pivots are the partial-pivoting choice
  const int m = n - K;   // panel rows
  for (size_t e = threadIdx.x; e < (size_t)m * b; e += blockDim.x) {
    const int i = (int)(e / b), jj = (int)(e % b);
    W[e] = A[(size_t)(K + i) * n + K + jj];
  }
  __syncthreads();
  for (int j = 0; j < b; ++j) {
    double best = -1.0;
    int bi = j;
    for (int i = j + threadIdx.x; i < m; i += blockDim.x) {
      const double v = fabs(W[(size_t)i * b + j]);
      if (v > best) { best = v; bi = i; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) {
        const double o = sv[threadIdx.x + st];
        const int oi = si[threadIdx.x + st];
        if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
          sv[threadIdx.x] = o;
          si[threadIdx.x] = oi;
        }
      }
      __syncthreads();
    }
    const int p = si[0];
    if (threadIdx.x == 0) {
      piv[K + j] = K + p;
      if (sv[0] == 0.0) singular[0] = 1;
    }
    if (p != j && threadIdx.x < b) {
      const double t = W[(size_t)j * b + threadIdx.x];
      W[(size_t)j * b + threadIdx.x] = W[(size_t)p * b + threadIdx.x];
      W[(size_t)p * b + threadIdx.x] = t;
    }
    __syncthreads();
    const double d = W[(size_t)j * b + j];
    const double id = (d != 0.0) ? 1.0 / d : 0.0;
    for (int i = j + 1 + threadIdx.x; i < m; i += blockDim.x) {
      double* wi = W + (size_t)i * b;
      const double l = wi[j] * id;
      wi[j] = l;
      for (int c = j + 1; c < b; ++c) wi[c] -= l * W[(size_t)j * b + c];
    }
    __syncthreads();
  }
}

// The same pivot choice with the panel rows in registers: thread t holds
// rows t, t + 512, ... (RPT rows, m <= 512 RPT) of six panel columns at a
// time (eight spill registers).  Per column: the pivot (largest magnitude, the lowest row on ties, as
// a sequential search) by a wave reduction and one over the 8 waves, the
// exchange of the two rows through LDS (their later panel columns are swapped
// in W), the elimination in registers.  After six columns the panel's later
// columns take one rank-6 update in W: U12 = L11^-1 A12 by one thread per
// column, then A22 -= L21 U12 by the row owners.  Only the pivots leave the
// kernel; two barriers per column instead of a pass over the panel in global
// memory.
constexpr int kPanelThreads = 512, kPanelRpt = 10, kPanelSub = 6;

template <int RPT>
__global__ __launch_bounds__(kPanelThreads) void k_bgj_panel_reg(int n, int K, int b, const double* __restrict__ A,
                                                                 double* __restrict__ W, int* __restrict__ piv,
                                                                 int* __restrict__ singular,
                                                                 const int* __restrict__ need) {
  constexpr int NT = kPanelThreads, SUB = kPanelSub, NW = NT / 64;
  __shared__ double s_best[NW];
  __shared__ int s_bi[NW];
  __shared__ double s_prow[SUB], s_jrow[SUB];
  __shared__ double s_l[SUB][SUB];
  __shared__ double s_u[SUB][kBgjBlock];
  if (!*need) return;   // the diagonal pivots are the partial-pivoting choice
  const int m = n - K, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (size_t e = tid; e < (size_t)m * b; e += NT) {
    const int i = (int)(e / b), jj = (int)(e % b);
    W[e] = A[(size_t)(K + i) * n + K + jj];
  }
  __syncthreads();
  for (int s0 = 0; s0 < b; s0 += SUB) {
    const int ws = b - s0 < SUB ? b - s0 : SUB;
    const int rest = b - s0 - SUB;   // later panel columns (> 0 only when ws == SUB)
    double v[RPT][SUB];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int i = tid + k * NT;
#pragma unroll
      for (int c = 0; c < SUB; ++c) v[k][c] = (i < m && c < ws) ? W[(size_t)i * b + s0 + c] : 0.0;
    }
#pragma unroll
    for (int jj = 0; jj < SUB; ++jj) {
      if (jj >= ws) break;   // uniform
      const int j = s0 + jj;
      double best = -1.0;
      int bi = 0x7fffffff;
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int i = tid + k * NT;
        if (i >= j && i < m) {
          const double a = fabs(v[k][jj]);
          if (a > best) { best = a; bi = i; }
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      if (lane == 0) { s_best[wv] = best; s_bi[wv] = bi; }
      __syncthreads();
      best = s_best[0];
      bi = s_bi[0];
      for (int w = 1; w < NW; ++w) {
        const double ob = s_best[w];
        const int oi = s_bi[w];
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      }
      // no row with a number (every candidate NaN): keep row j, report singular
      const bool none = bi == 0x7fffffff;
      const int p = none ? j : bi;
      const int tp = p % NT, kp = p / NT, tj = j % NT, kj = j / NT;
      if (tid == tp) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
          if (k == kp)
#pragma unroll
            for (int c = 0; c < SUB; ++c) s_prow[c] = v[k][c];
      }
      if (tid == tj) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
          if (k == kj)
#pragma unroll
            for (int c = 0; c < SUB; ++c) s_jrow[c] = v[k][c];
      }
      if (tid == 0) {
        piv[K + j] = K + p;
        if (none || !(best > 0.0)) singular[0] = 1;
      }
      if (p != j)
        for (int t = tid; t < rest; t += NT) {
          double* wj = W + (size_t)j * b + s0 + SUB + t;
          double* wp = W + (size_t)p * b + s0 + SUB + t;
          const double x = *wj;
          *wj = *wp;
          *wp = x;
        }
      __syncthreads();
      if (tid == tj) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
          if (k == kj)
#pragma unroll
            for (int c = 0; c < SUB; ++c) v[k][c] = s_prow[c];
      }
      if (tid == tp && p != j) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
          if (k == kp)
#pragma unroll
            for (int c = 0; c < SUB; ++c) v[k][c] = s_jrow[c];
      }
      const double d = s_prow[jj];
      const double id = (d != 0.0) ? 1.0 / d : 0.0;
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int i = tid + k * NT;
        if (i > j && i < m) {
          const double l = v[k][jj] * id;
          v[k][jj] = l;
#pragma unroll
          for (int c = jj + 1; c < SUB; ++c) v[k][c] -= l * s_prow[c];
        }
      }
    }
    if (rest > 0) {
      // L11 (unit lower) of the rows s0 .. s0+SUB-1
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int i = tid + k * NT;
        if (i >= s0 && i < s0 + SUB)
#pragma unroll
          for (int c = 0; c < SUB; ++c) s_l[i - s0][c] = v[k][c];
      }
      __syncthreads();
      for (int c = tid; c < rest; c += NT) {
        double u[SUB];
#pragma unroll
        for (int t = 0; t < SUB; ++t) {
          double x = W[(size_t)(s0 + t) * b + s0 + SUB + c];
#pragma unroll
          for (int q = 0; q < t; ++q) x -= s_l[t][q] * u[q];
          u[t] = x;
          s_u[t][c] = x;
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int i = tid + k * NT;
        if (i >= s0 + SUB && i < m) {
          double* wi = W + (size_t)i * b + s0 + SUB;
          for (int c = 0; c < rest; ++c) {
            double x = wi[c];
#pragma unroll
            for (int t = 0; t < SUB; ++t) x -= v[k][t] * s_u[t][c];
            wi[c] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// one thread per column: the swaps (K + j, piv[K + j]), j = 0 .. b-1, in order
__global__ void k_bgj_swap(int n, int K, int b, double* __restrict__ A, const int* __restrict__ piv) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  for (int j = 0; j < b; ++j) {
    const int r = piv[K + j];
    if (r != K + j) {
      const double t = A[(size_t)(K + j) * n + c];
      A[(size_t)(K + j) * n + c] = A[(size_t)r * n + c];
      A[(size_t)r * n + c] = t;
    }
  }
}

// P = A[K:K+b, K:K+b]^-1 by in-place Gauss-Jordan without pivoting, padded
// to bp x bp with zeros.  Per step the pivot row and column are copied first,
// so every entry is updated in one pass from the old values (two barriers);
// thread (c, g) updates column c of the rows g, g + 4, ..., all 16 reads of
// a step issued together.
//
// spec = 1 runs first, on the block as it stands (diagonal pivots): it marks
// the pivots as the identity, keeps row k of step k (the row of U in the LU
// of the block) for U, and sets *need when a pivot is zero or an entry below
// it is larger in magnitude -- then partial pivoting would swap rows and the
// panel kernel decides the pivots, after which a spec = 0 launch inverts the
// swapped block.  spec = 0 does nothing unless *need is set.
__global__ __launch_bounds__(256) void k_bgj_block_inverse(int n, int K, int b, int bp, const double* __restrict__ A,
                                                           double* __restrict__ P, double* __restrict__ U,
                                                           int* __restrict__ need, int* __restrict__ piv,
                                                           int* __restrict__ singular, int spec) {
  __shared__ double B[kBgjBlock][kBgjBlock + 1];
  __shared__ double Us[kBgjBlock][kBgjBlock];   // U rows, stored once at the end
  __shared__ double rowk[kBgjBlock], colk[kBgjBlock];
  __shared__ int s_need;
  if (!spec && !*need) return;
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;   // 4 row groups
  if (threadIdx.x == 0) s_need = 0;
  if (spec && (int)threadIdx.x < b) piv[K + threadIdx.x] = K + threadIdx.x;
#pragma unroll
  for (int m = 0; m < kBgjBlock / 4; ++m) {
    const int i = g + 4 * m;
    B[i][c] = (i < b && c < b) ? A[(size_t)(K + i) * n + K + c] : 0.0;
  }
  for (int k = 0; k < b; ++k) {
    __syncthreads();
    const double d = B[k][k];
    if (d == 0.0) {
      if (threadIdx.x == 0) {
        if (spec) *need = 1;
        else singular[0] = 1;
      }
      return;
    }
    if ((int)threadIdx.x < b) {
      rowk[threadIdx.x] = B[k][threadIdx.x];
      colk[threadIdx.x] = B[threadIdx.x][k];
      if (spec) {
        Us[k][threadIdx.x] = ((int)threadIdx.x >= k) ? B[k][threadIdx.x] : 0.0;
        if ((int)threadIdx.x > k && !(fabs(B[threadIdx.x][k]) <= fabs(d))) s_need = 1;
      }
    }
    __syncthreads();
    const double pk = 1.0 / d;
    const double rc = rowk[c] * pk;
    // row k: B[k][c] pk, B[k][k] = pk; rows i != k: B[i][c] -= B[i][k] (B[k][c] pk),
    // B[i][k] = -B[i][k] pk
    double bi[kBgjBlock / 4], ci[kBgjBlock / 4];
#pragma unroll
    for (int m = 0; m < kBgjBlock / 4; ++m) {
      bi[m] = B[g + 4 * m][c];
      ci[m] = colk[g + 4 * m];
    }
#pragma unroll
    for (int m = 0; m < kBgjBlock / 4; ++m) {
      const int i = g + 4 * m;
      if (i < b && c < b) {
        double x;
        if (i == k) x = (c == k) ? pk : rc;
        else if (c == k) x = -ci[m] * pk;
        else x = bi[m] - ci[m] * rc;
        B[i][c] = x;
      }
    }
  }
  __syncthreads();
  if (spec) {
    if (threadIdx.x == 0) *need = s_need;
    for (int e = threadIdx.x; e < b * kBgjBlock; e += blockDim.x) U[e] = Us[e >> 6][e & 63];
  }
  for (int e = threadIdx.x; e < bp * bp; e += blockDim.x) {
    const int t = e / bp, cc = e % bp;
    P[e] = (t < b && cc < b) ? B[t][cc] : 0.0;
  }
}

// Rows below the block under diagonal pivots: the multipliers of row i are
// l = a U^-1 (a = A[i][K:K+b], U from the speculative block inverse), and
// partial pivoting keeps the diagonal as long as every |a_j - sum_t<j l_t
// U_tj| <= |U_jj| (ties go to the lower row, the diagonal one).  One thread
// per row; sets *need otherwise.
__global__ __launch_bounds__(256) void k_bgj_check_below(int n, int K, int b, const double* __restrict__ A,
                                                         const double* __restrict__ U, int* __restrict__ need) {
  __shared__ double Us[kBgjBlock][kBgjBlock + 1];
  for (int e = threadIdx.x; e < kBgjBlock * kBgjBlock; e += blockDim.x) {
    const int t = e >> 6, j = e & 63;
    Us[t][j] = (t < b && j < b && j >= t) ? U[e] : 0.0;
  }
  __syncthreads();
  const int i = K + b + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* a = A + (size_t)i * n + K;
  double l[kBgjBlock];   // register array: constant indices after unrolling
  bool bad = false;
#pragma unroll
  for (int j = 0; j < kBgjBlock; ++j) {
    if (j < b) {
      double r = a[j];
#pragma unroll
      for (int t = 0; t < j; ++t) r -= l[t] * Us[t][j];
      bad |= !(fabs(r) <= fabs(Us[j][j]));
      l[j] = r / Us[j][j];
    } else {
      l[j] = 0.0;
    }
  }
  if (bad) *need = 1;
}

// Rows K..K+b-1 in place: A[K:K+b, c] = P A[K:K+b, c] outside the block
// columns and P inside them.  One workgroup per 64 columns: the 64 x 64
// operands in LDS, the product on FP64 matrix cores (wave w: rows
// 16w..16w+15 of the 64 x 64 result, v_mfma_f64_16x16x4f64 with A operand
// (m = lane & 15, k = lane >> 4), B operand (k = lane >> 4, n = lane & 15)
// and accumulator entry t at (m = 4t + (lane >> 4), n = lane & 15)).
__global__ __launch_bounds__(256) void k_bgj_rows(int n, int K, int b, int bp, double* __restrict__ A,
                                                  const double* __restrict__ P, double* __restrict__ R) {
  __shared__ double Ps[kBgjBlock][kBgjBlock + 1];
  __shared__ double Ak[kBgjBlock][kBgjBlock + 1];
  const int c0 = blockIdx.x * 64, cc = threadIdx.x & 63, g = threadIdx.x >> 6;
  for (int r = g; r < kBgjBlock; r += 4) {
    Ps[r][cc] = (r < b && cc < b) ? P[r * bp + cc] : 0.0;
    Ak[r][cc] = (r < b && c0 + cc < n) ? A[(size_t)(K + r) * n + c0 + cc] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v4d_bgj acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = v4d_bgj{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k0 = 0; k0 < kBgjBlock; k0 += 4) {
    const double av = Ps[16 * w + (lane & 15)][k0 + (lane >> 4)];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Ak[k0 + (lane >> 4)][16 * j + (lane & 15)], acc[j], 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int r = 16 * w + 4 * t + (lane >> 4), c = c0 + 16 * j + (lane & 15);
      if (c < n && r < bp) R[(size_t)r * n + c] = acc[j][t];   // zero rows r >= b (P padded with zeros)
      if (r < b && c < n) A[(size_t)(K + r) * n + c] = (c >= K && c < K + b) ? Ps[r][c - K] : acc[j][t];
    }
}

// C[i][s] = A[i][K + s] for the rows outside the block (zero rows inside it
// and zero columns s >= b up to bp): a compact copy of the operand of
// k_bgj_update, taken before k_bgj_cols overwrites the block columns
__global__ void k_bgj_make_c(int n, int K, int b, int bp, const double* __restrict__ A, double* __restrict__ C) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6), s = threadIdx.x & 63;
  if (i >= n || s >= bp) return;
  C[(size_t)i * bp + s] = (s < b && (i < K || i >= K + b)) ? A[(size_t)i * n + K + s] : 0.0;
}

// A[o, o'] -= C R for the rows o and columns o' outside the block (C = the
// old block columns, R = P A[K, :]; the block rows already hold R, the block
// columns are rewritten by k_bgj_cols), on 64 x 64 tiles with
// v_mfma_f64_16x16x4f64: wave w owns tile rows 16w .. 16w+15 and four
// 16-column MFMA tiles; operands straight from the (cache-resident) C and R.
// Column tiles from tile tj0 on; the tile column starting at skip_col (-1:
// none) is left out (the look-ahead strip, updated by its own launch).
__global__ __launch_bounds__(256) void k_bgj_update(int n, int K, int b, int bp, const double* __restrict__ C,
                                                    const double* __restrict__ R, double* __restrict__ A, int tj0,
                                                    int skip_col) {
  const int ti = blockIdx.y * 64, tj = (blockIdx.x + tj0) * 64;
  if (tj == skip_col) return;
  if ((ti >= K && ti + 64 <= K + b) || (tj >= K && tj + 64 <= K + b)) return;   // block rows / columns only
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int arow = ti + 16 * w + (lane & 15);
  v4d_bgj acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = v4d_bgj{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < bp; k0 += 4) {
    const int k = k0 + (lane >> 4);
    const double av = (arow < n) ? C[(size_t)arow * bp + k] : 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = tj + 16 * j + (lane & 15);
      const double bv = (col < n) ? R[(size_t)k * n + col] : 0.0;
      acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int row = ti + 16 * w + 4 * t + (lane >> 4), col = tj + 16 * j + (lane & 15);
      if (row < n && col < n && (row < K || row >= K + b) && (col < K || col >= K + b))
        A[(size_t)row * n + col] -= acc[j][t];
    }
}

// The block columns of the other rows in place: A[o, K:K+b] = -A[o, K:K+b] P,
// 64 rows per workgroup on FP64 matrix cores (layout as in k_bgj_rows).
__global__ __launch_bounds__(256) void k_bgj_cols(int n, int K, int b, int bp, double* __restrict__ A,
                                                  const double* __restrict__ P) {
  __shared__ double Ps[kBgjBlock][kBgjBlock + 1];
  __shared__ double Cs[kBgjBlock][kBgjBlock + 1];
  const int i0 = blockIdx.x * 64, cc = threadIdx.x & 63, g = threadIdx.x >> 6;
  for (int r = g; r < kBgjBlock; r += 4) {
    Ps[r][cc] = (r < b && cc < b) ? P[r * bp + cc] : 0.0;
    const int i = i0 + r;
    Cs[r][cc] = (i < n && cc < b) ? A[(size_t)i * n + K + cc] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v4d_bgj acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = v4d_bgj{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k0 = 0; k0 < kBgjBlock; k0 += 4) {
    const double av = Cs[16 * w + (lane & 15)][k0 + (lane >> 4)];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Ps[k0 + (lane >> 4)][16 * j + (lane & 15)], acc[j], 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = i0 + 16 * w + 4 * t + (lane >> 4), sidx = 16 * j + (lane & 15);
      if (i < n && (i < K || i >= K + b) && sidx < b) A[(size_t)i * n + K + sidx] = -acc[j][t];
    }
}

// out[r][c] = A[r][q[c]]
__global__ void k_bgj_unpermute_to(int n, const double* __restrict__ A, const int* __restrict__ q,
                                   double* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  out[e] = A[(size_t)r * n + q[c]];
}

// A[r][c] = A[r][q[c]] in place: one workgroup per row, the row staged in
// LDS (n doubles of dynamic shared memory, n <= 8192)
__global__ __launch_bounds__(256) void k_bgj_unpermute(int n, double* __restrict__ A, const int* __restrict__ q) {
  extern __shared__ double row[];
  double* a = A + (size_t)blockIdx.x * n;
  for (int c = threadIdx.x; c < n; c += blockDim.x) row[c] = a[c];
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) a[c] = row[q[c]];
}

template <typename T>
T* bgj_alloc(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw std::runtime_error("dense inverse: out of device memory");
  return static_cast<T*>(p);
}

}  // namespace

namespace {
// D[r][col] = val for the entries of row r (D zeroed before)
__global__ void k_csr_scatter_dense(int n, const int* __restrict__ rp, const int* __restrict__ col,
                                    const double* __restrict__ val, double* __restrict__ D) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int k = rp[r]; k < rp[r + 1]; ++k) D[(size_t)r * n + col[k]] = val[k];
}
}  // namespace

void csr_to_dense(const DevCSR& A, double* D, hipStream_t st) {
  const size_t nn = (size_t)A.n * A.n;
  if (nn == 0) return;
  hipLaunchKernelGGL(k_fill<double>, dim3(4096), dim3(256), 0, st, (int)nn, 0.0, D);   // n <= 20000
  hipLaunchKernelGGL(k_csr_scatter_dense, dim3((A.n + 255) / 256), dim3(256), 0, st, A.n, A.row_ptr, A.col, A.val, D);
}

bool dense_invert(int n, double* A, hipStream_t st) {
  if (n <= 0) return true;
  const int nb = kBgjBlock;
  double* W = bgj_alloc<double>((size_t)n * nb);
  double* C = bgj_alloc<double>((size_t)n * nb);
  double* R = bgj_alloc<double>((size_t)n * nb);
  double* P2 = bgj_alloc<double>((size_t)2 * nb * nb);   // block inverses of even / odd blocks
  // rows up to 8,192 doubles are unpermuted in place through LDS; longer ones
  // through a second matrix
  const bool in_place = (size_t)n * sizeof(double) <= 65536;
  double* T = in_place ? nullptr : bgj_alloc<double>((size_t)n * n);
  double* U = bgj_alloc<double>((size_t)nb * nb);
  int* piv = bgj_alloc<int>(n);
  int* q = bgj_alloc<int>(n);
  int* sing = bgj_alloc<int>(2);   // [0] singular, [1] pivoting needed for the current panel
  int* need = sing + 1;
  // look-ahead: while the trailing update of block K runs, a second stream
  // updates the next block's column strip and runs that block's speculative
  // inverse and check below it (they read nothing else the update writes)
  hipStream_t s2 = nullptr;
  hipEvent_t ev_rows = nullptr, ev_next = nullptr;
  if (hipStreamCreate(&s2) != hipSuccess || hipEventCreateWithFlags(&ev_rows, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ev_next, hipEventDisableTiming) != hipSuccess)
    throw std::runtime_error("dense inverse: stream or event creation failed");
  auto release = [&]() {
    (void)hipStreamSynchronize(s2);
    (void)hipEventDestroy(ev_rows); (void)hipEventDestroy(ev_next); (void)hipStreamDestroy(s2);
    (void)hipFree(W); (void)hipFree(C); (void)hipFree(R); (void)hipFree(P2); (void)hipFree(U);
    if (T) (void)hipFree(T);
    (void)hipFree(piv); (void)hipFree(q); (void)hipFree(sing);
  };
  (void)hipMemsetAsync(sing, 0, 2 * sizeof(int), st);
  const int g1 = (n + 255) / 256;
  const dim3 tiles((n + 63) / 64, (n + 63) / 64);
  // diagonal pivots first; the pivoting panel and a second block inverse
  // only when partial pivoting would swap rows in a panel
  auto speculate = [&](int K, hipStream_t s) {
    const int b = n - K < nb ? n - K : nb;
    const int bp = (b + 3) / 4 * 4;
    double* P = P2 + (size_t)((K / nb) & 1) * nb * nb;
    hipLaunchKernelGGL(k_bgj_block_inverse, dim3(1), dim3(256), 0, s, n, K, b, bp, A, P, U, need, piv, sing, 1);
    if (n - K - b > 0)
      hipLaunchKernelGGL(k_bgj_check_below, dim3((n - K - b + 255) / 256), dim3(256), 0, s, n, K, b, A, U, need);
  };
  speculate(0, st);
  for (int K = 0; K < n; K += nb) {
    const int b = n - K < nb ? n - K : nb;
    const int bp = (b + 3) / 4 * 4;
    double* P = P2 + (size_t)((K / nb) & 1) * nb * nb;
    if (K > 0) (void)hipStreamWaitEvent(st, ev_next, 0);   // this block's speculation (second stream)
    if (n - K <= kPanelThreads * kPanelRpt)
      hipLaunchKernelGGL(k_bgj_panel_reg<kPanelRpt>, dim3(1), dim3(kPanelThreads), 0, st, n, K, b, A, W, piv, sing,
                         need);
    else
      hipLaunchKernelGGL(k_bgj_panel, dim3(1), dim3(1024), 0, st, n, K, b, A, W, piv, sing, need);
    hipLaunchKernelGGL(k_bgj_swap, dim3(g1), dim3(256), 0, st, n, K, b, A, piv);
    hipLaunchKernelGGL(k_bgj_block_inverse, dim3(1), dim3(256), 0, st, n, K, b, bp, A, P, U, need, piv, sing, 0);
    hipLaunchKernelGGL(k_bgj_make_c, dim3((n + 3) / 4), dim3(256), 0, st, n, K, b, bp, A, C);
    hipLaunchKernelGGL(k_bgj_rows, dim3((n + 63) / 64), dim3(256), 0, st, n, K, b, bp, A, P, R);
    const int Kn = K + nb;
    if (Kn < n) {
      // the next block's column strip, then its speculation, on the second stream
      (void)hipEventRecord(ev_rows, st);
      (void)hipStreamWaitEvent(s2, ev_rows, 0);
      hipLaunchKernelGGL(k_bgj_update, dim3(1, tiles.y), dim3(256), 0, s2, n, K, b, bp, C, R, A, Kn / 64, -1);
      speculate(Kn, s2);
      (void)hipEventRecord(ev_next, s2);
    }
    hipLaunchKernelGGL(k_bgj_update, tiles, dim3(256), 0, st, n, K, b, bp, C, R, A, 0, Kn < n ? Kn : -1);
    hipLaunchKernelGGL(k_bgj_cols, dim3((n + 63) / 64), dim3(256), 0, st, n, K, b, bp, A, P);
  }
  // the swaps in order make the row permutation p (row i of the permuted
  // matrix = row p[i]); A^-1[r][p[k]] = M[r][k], i.e. column c comes from
  // column q[c] = p^-1[c]
  std::vector<int> hp(n), hq(n);
  int hs = 0;
  if (hipMemcpyAsync(hp.data(), piv, n * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(&hs, sing, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    release();
    throw std::runtime_error("dense inverse: device error");
  }
  if (hs) {
    release();
    return false;
  }
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int j = 0; j < n; ++j) std::swap(perm[j], perm[hp[j]]);
  for (int k = 0; k < n; ++k) hq[perm[k]] = k;
  (void)hipMemcpyAsync(q, hq.data(), n * sizeof(int), hipMemcpyHostToDevice, st);
  if (in_place) {
    hipLaunchKernelGGL(k_bgj_unpermute, dim3(n), dim3(256), (size_t)n * sizeof(double), st, n, A, q);
  } else {
    hipLaunchKernelGGL(k_bgj_unpermute_to, dim3((int)(((size_t)n * n + 255) / 256)), dim3(256), 0, st, n, A, q, T);
    (void)hipMemcpyAsync(A, T, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st);
  }
  const bool ok = hipStreamSynchronize(st) == hipSuccess;
  release();
  if (!ok) throw std::runtime_error("dense inverse: device error");
  return true;
}

}  // namespace pmx

// dense row-major y = A x (one 64-lane wavefront per row), x/y FP64 or FP32
namespace pmx {
namespace {
template <typename T, typename AT = double>
__global__ __launch_bounds__(256) void k_dense_mv(int n, const AT* __restrict__ A,
                                                  const T* __restrict__ x, T* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  const AT* a = A + (size_t)row * n;
  // four independent partial sums keep four row loads in flight per lane
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int j = lane;
  for (; j + 192 < n; j += 256) {
    const double a0 = __builtin_nontemporal_load(a + j), a1 = __builtin_nontemporal_load(a + j + 64);
    const double a2 = __builtin_nontemporal_load(a + j + 128), a3 = __builtin_nontemporal_load(a + j + 192);
    s0 += (double)a0 * (double)x[j];
    s1 += (double)a1 * (double)x[j + 64];
    s2 += (double)a2 * (double)x[j + 128];
    s3 += (double)a3 * (double)x[j + 192];
  }
  for (; j < n; j += 64) s0 += (double)a[j] * (double)x[j];
  double s = group_sum<64>((s0 + s1) + (s2 + s3));
  if (lane == 0) y[row] = (T)s;
}
}  // namespace

void dense_mv(int n, const double* A, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_dense_mv<double>, dim3((n + 3) / 4), dim3(256), 0, st, n, A, x, y);
}
void dense_mv_f32(int n, const double* A, const float* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_dense_mv<float>, dim3((n + 3) / 4), dim3(256), 0, st, n, A, x, y);
}
void dense_mv_f32a(int n, const float* A, const float* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL((k_dense_mv<float, float>), dim3((n + 3) / 4), dim3(256), 0, st, n, A, x, y);
}
}  // namespace pmx
