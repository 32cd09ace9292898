// Deterministic workgroup reductions shared by the solver translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lfgpu {

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// Sum over a workgroup of BLOCK threads (fixed shape, fixed order); the
// result is valid in thread 0.  sh needs BLOCK / 64 doubles.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < BLOCK / 64; ++q) t += sh[q];
  }
  return t;
}

}  // namespace lfgpu
