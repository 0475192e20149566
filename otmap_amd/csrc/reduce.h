// Deterministic two-pass reductions shared by the HIP kernels.
//
// Pass 1: every workgroup reduces its values in a fixed order and writes one
// partial per value (partials[k * nb + block]).  Pass 2: a single-workgroup
// scalar kernel, launched next on the same stream, folds the partials in a
// fixed order and runs the scalar logic of the solver (beta, the quartic line
// search, convergence) on the device.  The kernel boundary orders the two
// passes, so no atomics or fences are needed, and results do not depend on
// scheduling.  Requires NT threads per workgroup.
#pragma once

#include "otm_internal.h"

namespace otm {

__device__ static inline int block_id() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ static inline int block_count() { return gridDim.x * gridDim.y; }

// v moved across lanes by one DPP pattern (CTRL) in the rows of ROWS; lanes
// without a source (and rows outside ROWS) get 0.  Two 32-bit moves.
template <int CTRL, int ROWS>
__device__ static inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Wave sums of K values on the VALU (no LDS traffic): row shifts by 1, 2, 4,
// 8 leave each 16-lane row's sum in its lane 15, the row broadcasts 15 and 31
// add rows 0 + 1 into lane 31 and rows 0 + 1 + 2 + 3 into lane 63.  Lane 63
// holds the totals.  Fixed order; every step moves all K values at once.
template <int K>
__device__ static inline void wave_sum(double (&v)[K]) {
#define OTM_DPP_STEP(CTRL, ROWS)                                  \
  {                                                               \
    double y[K];                                                  \
    _Pragma("unroll") for (int k = 0; k < K; ++k) y[k] = dpp_move<CTRL, ROWS>(v[k]); \
    _Pragma("unroll") for (int k = 0; k < K; ++k) v[k] += y[k];   \
  }
  OTM_DPP_STEP(0x111, 0xf)  // row_shr:1
  OTM_DPP_STEP(0x112, 0xf)  // row_shr:2
  OTM_DPP_STEP(0x114, 0xf)  // row_shr:4
  OTM_DPP_STEP(0x118, 0xf)  // row_shr:8
  OTM_DPP_STEP(0x142, 0xa)  // row_bcast:15 into rows 1 and 3
  OTM_DPP_STEP(0x143, 0xc)  // row_bcast:31 into rows 2 and 3
#undef OTM_DPP_STEP
}

// In-block sum of K values over NTH threads (a multiple of 64); the result
// is valid in thread 0.  Fixed order.
template <int K, int NTH = NT>
__device__ static inline void block_sum(double (&v)[K]) {
  constexpr int NW = NTH / 64;
  __shared__ double red[K][NW];
  wave_sum<K>(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 63)
    for (int k = 0; k < K; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) {
      double x = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
      for (int q = 4; q < NW; q += 4) x += (red[k][q] + red[k][q + 1]) + (red[k][q + 2] + red[k][q + 3]);
      v[k] = x;
    }
}

// pass 1: K workgroup partials of workgroup bid among nb (NTH threads per workgroup)
template <int K, int NTH = NT>
__device__ static inline void write_partials_at(double (&v)[K], double* partials, int bid, int nb) {
  static_assert(NTH % 256 == 0, "workgroups of 4k waves");
  block_sum<K, NTH>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) gp(partials)[long(k) * nb + bid] = v[k];
}

template <int K>
__device__ static inline void write_partials(double (&v)[K], double* partials) {
  write_partials_at<K>(v, partials, block_id(), block_count());
}

// XCD-aware tile order.  Workgroups are dealt to the 8 XCDs round-robin
// (workgroup b runs on XCD b % 8), each with its own L2.  When the R x C tile
// grid splits into 4 x 2 blocks, local tile lt goes to block lt % 8, so one
// XCD's workgroups share R/4 tile rows of A and C/2 tile columns of B in its
// L2 instead of streaming the whole operands.  Requires the job's first
// workgroup index to be a multiple of 8 (job tables pad GEMM launches).
__device__ static inline void xcd_tile(int lt, int R, int C, int& tm, int& tn) {
  if (R % 4 == 0 && C % 2 == 0) {
    const int xcd = lt & 7, local = lt >> 3;
    const int br = R >> 2, bc = C >> 1;
    tm = (xcd >> 1) * br + local / bc;
    tn = (xcd & 1) * bc + local % bc;
  } else {
    tm = lt / C;
    tn = lt % C;
  }
}

// Job of workgroup b in a launch over a job table: the kind's block holds
// count + 1 prefix sums of the workgroups per job, then the job of every
// workgroup (one load instead of a search); uniform across the workgroup.
__device__ static inline int job_of(const int* __restrict__ starts, int count, int b) {
  return starts[count + 1 + b];
}

// pass 2 (single workgroup): fold nb partials of each of K values; valid in
// thread 0.  Each thread issues the FOLD_BATCH loads of all K values before
// adding (one latency per batch, not one per value).
constexpr int FOLD_BATCH = 8;

// this thread's share of the fold (added to s; block_sum completes it)
template <int K>
__device__ static inline void fold_share(double (&s)[K], const double* partials, int nb) {
  for (int base = 0; base < nb; base += NT * FOLD_BATCH) {
    double tmp[K][FOLD_BATCH];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int u = 0; u < FOLD_BATCH; ++u) {
        const int j = base + u * NT + threadIdx.x;
        tmp[k][u] = (j < nb) ? gp(partials)[long(k) * nb + j] : 0.;
      }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int u = 0; u < FOLD_BATCH; ++u) s[k] += tmp[k][u];
  }
}

template <int K>
__device__ static inline void fold_partials(double (&v)[K], const double* partials, int nb) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.;
  fold_share<K>(v, partials, nb);
  block_sum<K>(v);
}

// Wide pass 2: a workgroup of G groups of NT threads.  Group g folds
// batches g, g + G, ... (a batch is NT * FOLD_BATCH partials) in
// fold_share's order, so more than one batch costs fewer memory round trips;
// group_sum adds the group totals in group order.  With nb <= NT *
// FOLD_BATCH only group 0 reads partials and the others add exact zeros: the
// result equals the NT-thread fold's bit for bit.
template <int K, int G>
__device__ static inline void fold_share_grouped(double (&s)[K], const double* partials, int nb) {
  const int g = threadIdx.x / NT, lt = threadIdx.x % NT;
  for (int base = g * NT * FOLD_BATCH; base < nb; base += G * NT * FOLD_BATCH) {
    double tmp[K][FOLD_BATCH];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int u = 0; u < FOLD_BATCH; ++u) {
        const int j = base + u * NT + lt;
        tmp[k][u] = (j < nb) ? gp(partials)[long(k) * nb + j] : 0.;
      }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int u = 0; u < FOLD_BATCH; ++u) s[k] += tmp[k][u];
  }
}

// block_sum over G * NT threads: each group's four wave totals as block_sum
// adds them, then the groups in order; valid in thread 0.
template <int K, int G>
__device__ static inline void group_sum(double (&v)[K]) {
  static_assert(NT == 256, "four waves per group");
  constexpr int NW = G * NT / 64;
  __shared__ double red[K][NW];
  wave_sum<K>(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 63)
    for (int k = 0; k < K; ++k) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) {
      double x = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
      for (int g = 1; g < G; ++g) x += (red[k][4 * g] + red[k][4 * g + 1]) + (red[k][4 * g + 2] + red[k][4 * g + 3]);
      v[k] = x;
    }
}

}  // namespace otm
