// BF16 planes of the mixed-precision Laplacian solve: the 3-term split of a
// value and the blocked layout of a plane in HBM.  Shared by the transform
// kernels and the update pass that writes the next right-hand side.
#pragma once

#include <hip/hip_runtime.h>

namespace otm {

typedef unsigned short u16;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

__device__ inline u16 bf16_bits(float x) { return __builtin_bit_cast(u16, (__bf16)x); }
__device__ inline float bf16_float(u16 b) { return (float)__builtin_bit_cast(__bf16, b); }

// Element offset of (r, k) in a BF16 plane of row length K (K % 32 == 0)
// stored in blocks of 16 rows x 32 k (row blocks outer, k blocks inner); in a
// block, lane position l = 16 (k / 8 % 4) + r % 16 holds the 8 values
// k / 8 * 8 .. + 7 of row r, which is what lane l of v_mfma_f32_16x16x32_bf16
// takes as its A or B fragment.
__host__ __device__ inline long bofs(long r, long k, long K) {
  return ((r >> 4) * (K >> 5) + (k >> 5)) * 512 + ((((k >> 3) & 3) << 4) + (r & 15)) * 8 + (k & 7);
}

// exact 3-term split of an FP32 value
__device__ inline void split3(float x, u16& h0, u16& h1, u16& h2) {
  h0 = bf16_bits(x);
  const float r1 = x - bf16_float(h0);
  h1 = bf16_bits(r1);
  h2 = bf16_bits(r1 - bf16_float(h1));
}

// 3-term split of an FP64 value (24 significant bits)
__device__ inline void split3d(double x, u16& h0, u16& h1, u16& h2) {
  h0 = bf16_bits(float(x));
  const double r1 = x - double(bf16_float(h0));
  h1 = bf16_bits(float(r1));
  h2 = bf16_bits(float(r1 - double(bf16_float(h1))));
}

}  // namespace otm
