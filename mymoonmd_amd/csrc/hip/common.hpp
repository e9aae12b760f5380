// Shared device helpers for the CDNA4 (gfx950) kernels: 64-wide wavefront
// reductions and launch-shape constants.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pmx {

constexpr int kWave = 64;

// Sum over a group of G consecutive lanes (G a power of two <= 64) with
// cross-lane butterflies; every lane of the group gets the total.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// Block-wide sum of one value per thread; result valid in thread 0.
// Block size must be a multiple of 64 and at most 1024.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* lds /* >= 16 entries */) {
  v = group_sum<kWave>(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T s = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + kWave - 1) / kWave;
    for (int w = 0; w < nw; ++w) s += lds[w];
  }
  __syncthreads();
  return s;
}

__device__ __forceinline__ float bf16_to_float(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

// synchronises the lanes of one wavefront only (for waves that share no LDS
// with other waves): the fences order its LDS stores before the other lanes'
// loads, the barrier keeps the compiler from moving accesses across it
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace pmx
