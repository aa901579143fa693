// Deterministic FP64 grid reductions shared by the PCG, line-search and learned-correction kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace bd {

__device__ __forceinline__ void store_agent(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __double_as_longlong(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p) {
  return __longlong_as_double(__hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Grid-wide deterministic sum of NV doubles per thread. Returns true in thread 0 of the last block
// to finish, with the totals in tot. Every sum runs in a fixed order (per block: wave shuffles, then
// waves in order; the last block: thread t sums partials t, t + blockDim, ..., then the same block
// tree), so the result does not depend on which block finishes last.
// Ordering without fences: a block's partials are agent-scope atomic stores, and its thread waits for
// their completion (s_waitcnt on every counter) before the ticket's agent-scope increment, so a block
// that draws the last ticket finds every partial in place and reads them with agent-scope atomic
// loads. Nothing else passes between the blocks. An agent-scope fence would write back and then
// invalidate the whole L2 of the block's XCD, once per block (about 4 us per reduction of 1M nodes,
// more after kernels that leave the L2 dirty).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (&sh)[NV][16]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) sh[q][wid] = x;
  }
}

template <int NV>
__device__ bool grid_sum(double (&v)[NV], double* partials, unsigned int* ticket, double (&tot)[NV]) {
  __shared__ double sh[NV][16];
  __shared__ int last;
  const int nw = (blockDim.x + 63) >> 6;
  block_sum<NV>(v, sh);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < NV; q++) {
      double s = 0;
      for (int w = 0; w < nw; w++) s += sh[q][w];
      store_agent(&partials[blockIdx.x * NV + q], s);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const unsigned int t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) acc[q] = 0;
  for (unsigned int b = threadIdx.x; b < gridDim.x; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] += load_agent(&partials[b * NV + q]);
  block_sum<NV>(acc, sh);
  __syncthreads();
  if (threadIdx.x != 0) return false;
  for (int q = 0; q < NV; q++) {
    double s = 0;
    for (int w = 0; w < nw; w++) s += sh[q][w];
    tot[q] = s;
  }
  __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

}  // namespace bd
