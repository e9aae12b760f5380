// Snapshot / POD kernels (method of snapshots, src/ReducedOrderModels/POD.C).
//
// Snapshot storage: a column-major matrix over "global rows" with leading
// dimension ld: entry (g, j) = snapshot j of row g sits at S[g + j * ld].  A
// single problem is rows [0, n) with ld = n; a batch of problems stacks their
// DOFs (problem p owns rows [row0_p, row0_p + n_p)), so storing the current
// solution of every problem is one contiguous copy into column j.
//
// The Gram (autocorrelation) matrix G_p = S_p^T W S_p of every problem runs
// on the FP64 matrix cores (v_mfma_f64_16x16x4f64): the rows of a problem are
// cut into chunks of at most kPodChunk rows, one workgroup per chunk computes
// the chunk's m x m partial product, and a second kernel adds the partials of
// each problem in chunk order (deterministic).
#include "hip/common.hpp"
#include "hip/launch.hpp"

namespace pmx {

namespace {

typedef double v4d_pod __attribute__((ext_vector_type(4)));

// mean over the snapshots of every row, (1/m) * (((0 + s_0) + s_1) + ...),
// as POD::decompose_snaps (src/ReducedOrderModels/POD.C:534-561); with
// fluct the mean is subtracted from every snapshot
__global__ void k_pod_mean(int rows, int m, long long ld, double* __restrict__ S, double* __restrict__ mean,
                           int fluct) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows) return;
  double s = 0.0;
  for (int j = 0; j < m; ++j) s += S[g + j * ld];
  const double mu = (1.0 / m) * s;
  mean[g] = mu;
  if (fluct)
    for (int j = 0; j < m; ++j) S[g + j * ld] -= mu;
}

// WS = W S for the L2 inner product: W is a CSR matrix over the same rows
// (block diagonal for a batch, global column indices); one thread per row,
// each snapshot's row sum in column order
__global__ void k_pod_spmm(int rows, int m, long long ld, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ va, const double* __restrict__ S, double* __restrict__ WS) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows) return;
  const int s = rp[g], e = rp[g + 1];
  for (int j = 0; j < m; ++j) {
    double acc = 0.0;
    for (int k = s; k < e; ++k) acc += va[k] * S[ci[k] + j * ld];
    WS[g + j * ld] = acc;
  }
}

// partial Gram matrix of one chunk: P = S[r0:r0+rows, :]^T WS[r0:r0+rows, :]
// (m x m, row-major).  Four waves take a quarter of the rows each; a wave
// keeps a 64 x 64 block of P in 16 MFMA accumulators (tiles 16 x 16) and
// consumes 4 rows per MFMA step: lane l supplies S(row + l/16, a0 + l%16) as
// the A operand (the transposed snapshot tile) and WS(row + l/16, b0 + l%16)
// as the B operand; element t of lane l's result is P(a0 + 4t + l/16,
// b0 + l%16).  The waves' blocks are added in LDS in wave order.
__global__ __launch_bounds__(256) void k_pod_gram_partial(const int* __restrict__ chunk_row0,
                                                          const int* __restrict__ chunk_rows, int m, long long ld,
                                                          const double* __restrict__ S,
                                                          const double* __restrict__ WS,
                                                          double* __restrict__ partial) {
  __shared__ double blk[64 * 64];
  const int c = blockIdx.x;
  const int r0 = chunk_row0[c], nr = chunk_rows[c];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int per = (nr + 3) / 4;
  const int wr0 = r0 + wv * per, wr1 = min(r0 + nr, wr0 + per);
  double* out = partial + (size_t)c * m * m;
  const int nb = (m + 63) / 64;
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj < nb; ++bj) {
      const int a0 = 64 * bi, b0 = 64 * bj;
      const int nti = min(4, (m - a0 + 15) / 16), ntj = min(4, (m - b0 + 15) / 16);
      v4d_pod acc[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v4d_pod{0.0, 0.0, 0.0, 0.0};
      for (int r = wr0; r < wr1; r += 4) {
        const int row = r + (lane >> 4);
        const bool in = row < wr1;
        double av[4], bv[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int ca = a0 + 16 * t + (lane & 15), cb = b0 + 16 * t + (lane & 15);
          av[t] = (in && t < nti && ca < m) ? S[row + ca * ld] : 0.0;
          bv[t] = (in && t < ntj && cb < m) ? WS[row + cb * ld] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (i < nti && j < ntj) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
      }
      for (int w = 0; w < 4; ++w) {
        if (wv == w) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                const int a = 16 * i + 4 * t + (lane >> 4), b = 16 * j + (lane & 15);
                if (w == 0) blk[a * 64 + b] = acc[i][j][t];
                else blk[a * 64 + b] += acc[i][j][t];
              }
        }
        __syncthreads();
      }
      for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int a = a0 + e / 64, b = b0 + e % 64;
        if (a < m && b < m) out[(size_t)a * m + b] = blk[e];
      }
      __syncthreads();
    }
}

// G_p = sum of the partials of problem p's chunks, in chunk order
__global__ void k_pod_gram_reduce(int n_jobs, int m, const int* __restrict__ job_chunk0,
                                  const double* __restrict__ partial, double* __restrict__ G) {
  const long long mm = (long long)m * m;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_jobs * mm) return;
  const int p = (int)(t / mm);
  const long long e = t % mm;
  double s = 0.0;
  for (int c = job_chunk0[p]; c < job_chunk0[p + 1]; ++c) s += partial[c * mm + e];
  G[t] = s;
}

// basis: Phi(g, k) = sum_j S(g, j) C_p(j, k) for the problem p owning row g,
// C_p = eigenvectors / sqrt(eigenvalues) (m x kmax, column-major at
// C + p * m * kmax); the sum runs over j in order, as ublas::axpy_prod of a
// row-major matrix and a vector (src/ReducedOrderModels/POD.C:208-221)
__global__ void k_pod_basis(int rows, int n_jobs, const int* __restrict__ job_row0, const int* __restrict__ job_rank,
                            int m, int kmax, long long ld, const double* __restrict__ S,
                            const double* __restrict__ C, double* __restrict__ Phi) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows) return;
  int lo = 0, hi = n_jobs - 1;   // last job with job_row0 <= g
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (job_row0[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const double* Cp = C + (size_t)lo * m * kmax;
  for (int k = 0; k < job_rank[lo]; ++k) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += S[g + j * ld] * Cp[j + (size_t)k * m];
    Phi[g + k * ld] = acc;
  }
}

}  // namespace

void pod_mean(int rows, int m, long long ld, double* S, double* mean, bool fluct, hipStream_t st) {
  if (rows > 0)
    hipLaunchKernelGGL(k_pod_mean, dim3((rows + 255) / 256), dim3(256), 0, st, rows, m, ld, S, mean, fluct ? 1 : 0);
}

void pod_spmm(int rows, int m, long long ld, const int* rp, const int* ci, const double* va, const double* S,
              double* WS, hipStream_t st) {
  if (rows > 0)
    hipLaunchKernelGGL(k_pod_spmm, dim3((rows + 255) / 256), dim3(256), 0, st, rows, m, ld, rp, ci, va, S, WS);
}

void pod_gram(int n_chunks, const int* chunk_row0, const int* chunk_rows, int n_jobs, const int* job_chunk0, int m,
              long long ld, const double* S, const double* WS, double* partial, double* G, hipStream_t st) {
  if (n_chunks > 0)
    hipLaunchKernelGGL(k_pod_gram_partial, dim3(n_chunks), dim3(256), 0, st, chunk_row0, chunk_rows, m, ld, S, WS,
                       partial);
  const long long total = (long long)n_jobs * m * m;
  if (total > 0)
    hipLaunchKernelGGL(k_pod_gram_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, n_jobs, m,
                       job_chunk0, partial, G);
}

void pod_basis(int rows, int n_jobs, const int* job_row0, const int* job_rank, int m, int kmax, long long ld,
               const double* S, const double* C, double* Phi, hipStream_t st) {
  if (rows > 0)
    hipLaunchKernelGGL(k_pod_basis, dim3((rows + 255) / 256), dim3(256), 0, st, rows, n_jobs, job_row0, job_rank, m,
                       kmax, ld, S, C, Phi);
}

}  // namespace pmx
