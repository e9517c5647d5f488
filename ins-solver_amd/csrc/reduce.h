// Deterministic block reductions shared by the vector and SIPG kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace insp {

constexpr int kRedBlocks = 8192;  // partials per reduction slot (vector kernels)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a block of BLOCK threads; every thread returns the total. Fixed order.
template <int BLOCK>
__device__ __forceinline__ double block_sum1(double v) {
  __shared__ double sh[BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < BLOCK / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}

}  // namespace insp
