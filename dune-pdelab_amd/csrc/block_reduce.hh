// block_reduce.hh -- deterministic workgroup sums for the gfx950 kernels.
//
// Grid-wide sums are finished by a second, single-workgroup launch: an in-launch
// "last workgroup" reduction needs an agent-scope release fence per workgroup, which on
// this multi-XCD GPU writes back the whole L2 and costs far more than the launch.
#ifndef PDELAB_BLOCK_REDUCE_HH
#define PDELAB_BLOCK_REDUCE_HH

#include <hip/hip_runtime.h>

namespace pdelab_amd {

// Grid of the reducing kernels: kRedBlocks workgroups of kRedThreads write one partial
// (or NR partials) each, summed in order by one workgroup.
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

// Sum over the 64 lanes of a wave (valid in lane 0), in a fixed order.
template <typename A>
__device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Workgroup sum of NR values per thread through `scratch`, an LDS array of at least
// NR * WAVES values that the workgroup no longer reads (the kernel's own working array):
// no extra LDS, two barriers. The sums are valid in thread 0, in a fixed order.
template <int WAVES, typename A, int NR, typename S>
__device__ __forceinline__ void block_sum_reuse(A (&v)[NR], S* scratch) {
  A w[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) w[j] = wave_sum(v[j]);
  __syncthreads();  // every wave is done with the scratch array
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int j = 0; j < NR; ++j) scratch[j * WAVES + (threadIdx.x >> 6)] = (S)w[j];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      A s = A(0);
#pragma unroll
      for (int i = 0; i < WAVES; ++i) s += (A)scratch[j * WAVES + i];
      v[j] = s;
    }
}

// Sum over the workgroup of NR values per thread (THREADS <= 1024, any multiple of
// 32); the sums are valid in thread 0. Padded to whole waves with zeros in LDS, then
// reduced by wave 0 in a fixed order.
template <int THREADS, typename A, int NR>
__device__ __forceinline__ void block_sum_n(A (&v)[NR]) {
  constexpr int R = (THREADS + 63) / 64 * 64;
  __shared__ A red[NR][R];
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    red[j][t] = v[j];
    if (t < R - THREADS) red[j][THREADS + t] = A(0);
  }
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      A s = A(0);
#pragma unroll
      for (int m = 0; m < R / 64; ++m) s += red[j][t + 64 * m];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
      v[j] = s;
    }
  }
}

}  // namespace pdelab_amd

#endif
