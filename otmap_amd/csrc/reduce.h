// Deterministic workgroup reductions shared by the HIP kernels.
//
// A reduction writes one partial per workgroup (agent-scope stores), takes an
// atomic ticket, and the workgroup that draws the last ticket folds all
// partials in a fixed order and resets the ticket.  Results therefore do not
// depend on scheduling, and the follow-up scalar logic runs in that last
// workgroup instead of a separate launch.  Requires NT threads per workgroup.
#pragma once

#include "otm_internal.h"

namespace otm {

__device__ static inline int block_id() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ static inline int block_count() { return gridDim.x * gridDim.y; }

__device__ static inline void store_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ static inline double load_agent(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// In-block sum of K values; the result is valid in thread 0.  Fixed order.
template <int K>
__device__ static inline void block_sum(double (&v)[K]) {
  __shared__ double red[K][NT / 64];
  for (int k = 0; k < K; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < K; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) v[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
}

// Publishes K block partials (partials[k*nb + bid]) and returns true in the
// single workgroup that finishes last; that workgroup then owns the partials.
template <int K>
__device__ static inline bool publish_partials(double (&v)[K], double* partials, unsigned* counter) {
  __shared__ int is_last;
  block_sum<K>(v);
  const int nb = block_count();
  if (threadIdx.x == 0) {
    const int bid = block_id();
    for (int k = 0; k < K; ++k) store_agent(&partials[long(k) * nb + bid], v[k]);
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    is_last = (prev == unsigned(nb - 1)) ? 1 : 0;
    if (is_last) __threadfence();
  }
  __syncthreads();
  return is_last != 0;
}

// In the last workgroup: fold the K partial columns in a fixed order, reset
// the ticket.  Result valid in thread 0.
// Each thread issues FOLD_BATCH independent loads before adding them, so the
// latency of the agent-scope loads overlaps (a serial chain of such loads
// costs ~1 us each); the summation order stays fixed.
constexpr int FOLD_BATCH = 8;

template <int K>
__device__ static inline void fold_partials(double (&v)[K], const double* partials, unsigned* counter, int nb) {
  for (int k = 0; k < K; ++k) {
    double s = 0.;
    for (int base = 0; base < nb; base += NT * FOLD_BATCH) {
      double tmp[FOLD_BATCH];
#pragma unroll
      for (int u = 0; u < FOLD_BATCH; ++u) {
        const int j = base + u * NT + threadIdx.x;
        tmp[u] = (j < nb) ? load_agent(&partials[long(k) * nb + j]) : 0.;
      }
#pragma unroll
      for (int u = 0; u < FOLD_BATCH; ++u) s += tmp[u];
    }
    v[k] = s;
  }
  block_sum<K>(v);
  if (threadIdx.x == 0) *counter = 0u;
}

}  // namespace otm
