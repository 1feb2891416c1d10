// Small device helpers shared by the HIP sources (64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cytom {

// Exclusive scan of one value per thread across a block of BLOCK threads.
// Returns the thread's prefix; *total receives the block sum.  Every thread
// of the block must call it.
template <int BLOCK, typename T>
__device__ inline T block_scan_excl(T v, T* total) {
  static_assert(BLOCK % 64 == 0 && BLOCK <= 1024, "block size");
  __shared__ T wsum[BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  T woff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; ++w) {
    T s = wsum[w];
    if (w < wave) woff += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return woff + x - v;
}

// Sum across the block; the result is valid in every thread.
template <int BLOCK, typename T>
__device__ inline T block_sum(T v) {
  T tot;
  block_scan_excl<BLOCK, T>(v, &tot);
  return tot;
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace cytom
