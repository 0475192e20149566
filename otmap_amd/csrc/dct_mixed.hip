// Mixed-precision Laplacian solve: the same 2-D DCT-II diagonalisation as the
// FP64 path, with every GEMM run as a 3-term BF16 split on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), followed by FP64 iterative
// refinement against the exact operator.
//
// Split: x = x0 + x1 + x2 with x0 = bf16(x), x1 = bf16(x - x0), x2 =
// bf16(x - x0 - x1) (24 significant bits, exact for FP32 values).  A product
// keeps the six terms a_i b_j with i + j <= 2.
//
// Every stage is D[m][n] = sum_k A[m][k] B[n][k] with both operands stored
// with k contiguous (what the MFMA fragments want) and writes its output
// transposed through LDS, so the next stage again contracts its contiguous
// index:
//   stage 1  T1 = R C^T          A = R  [i][j]    B = C  [l][j]   -> T1^T [l][i]
//   stage 2  Y  = (C T1) / lam   A = T1^T[l][i]   B = C  [k][i]   -> Y    [k][l]
//   stage 3  U  = Y C            A = Y  [k][l]    B = C^T[j][l]   -> U^T  [j][k]
//   stage 4  X  = C^T U          A = U^T[j][k]    B = C^T[i][k]   -> X    [i][j] (FP64)
//
// Refinement (FP64): s = r' - M0 x with the exact stencil of the reference's
// operator (weight -1/2 at the four clamped diagonal neighbours, 2 on the
// diagonal; otlib/otsolver_2dgrid.cpp:284-306), then x += solve(s).

#include <cfloat>
#include <cmath>
#include <vector>

#include "otm_internal.h"
#include "reduce.h"

namespace otm {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int MB = 64, NB = 64, KB = 32;  // workgroup tile and K step
constexpr int LDK = KB + 8;               // LDS row length (bf16) of operand tiles
constexpr int LDD = MB + 4;               // LDS row length (f32) of the transposed output tile

enum MixedMode { X_PLAIN = 0, X_SPECTRAL = 1, X_FINAL = 2, X_FINAL_ADD = 3 };

__device__ inline u16 bf16_bits(float x) {
  __bf16 h = (__bf16)x;
  return __builtin_bit_cast(u16, h);
}
__device__ inline float bf16_float(u16 b) { return (float)__builtin_bit_cast(__bf16, b); }

// 3-term split of an FP32 value (exact)
__device__ inline void split3(float x, u16& h0, u16& h1, u16& h2) {
  h0 = bf16_bits(x);
  const float r1 = x - bf16_float(h0);
  h1 = bf16_bits(r1);
  const float r2 = r1 - bf16_float(h1);
  h2 = bf16_bits(r2);
}

// 3-term split of an FP64 value (relative residual ~2^-24)
__device__ inline void split3d(double x, u16& h0, u16& h1, u16& h2) {
  const float f0 = (float)x;
  h0 = bf16_bits(f0);
  const double r1 = x - (double)bf16_float(h0);
  h1 = bf16_bits((float)r1);
  const double r2 = r1 - (double)bf16_float(h1);
  h2 = bf16_bits((float)r2);
}

// D = A * B^T on 3-plane BF16 operands (plane stride `ps` elements), n x n.
// NP = number of split products (6: i+j<=2, 3: i+j<=1, 1: i=j=0).
template <int MODE, int NP>
__global__ __launch_bounds__(NT) void k_gemm_bf16x3(const u16* __restrict__ A, const u16* __restrict__ B, int n,
                                                   long ps, u16* __restrict__ out16, double* __restrict__ out64,
                                                   const double* __restrict__ add64, const double* __restrict__ sq,
                                                   SolveState* st, double* partials, unsigned* counter) {
  if (st && st->done) return;
  __shared__ __attribute__((aligned(16))) u16 As[2][3][MB * LDK];
  __shared__ __attribute__((aligned(16))) u16 Bs[2][3][NB * LDK];
  __shared__ float Ds[NB * LDD];
  const int m0 = blockIdx.y * MB, n0 = blockIdx.x * NB;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lr = lane & 15, lq = lane >> 4;
  const int lrow = t >> 2, lk = (t & 3) * 8;  // global->LDS: row, 8-element k chunk
  constexpr int NPL = (NP == 1) ? 1 : (NP == 3 ? 2 : 3);  // planes actually used
  uint4 ra[3], rb[3];
  auto gload = [&](int k0) {
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
      ra[p] = *reinterpret_cast<const uint4*>(A + p * ps + long(m0 + lrow) * n + k0 + lk);
      rb[p] = *reinterpret_cast<const uint4*>(B + p * ps + long(n0 + lrow) * n + k0 + lk);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
      *reinterpret_cast<uint4*>(&As[buf][p][lrow * LDK + lk]) = ra[p];
      *reinterpret_cast<uint4*>(&Bs[buf][p][lrow * LDK + lk]) = rb[p];
    }
  };
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < n; k0 += KB) {
    const bool more = k0 + KB < n;
    if (more) gload(k0 + KB);
    bf8 fa[2][3], fb[2][3];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int p = 0; p < NPL; ++p) {
        fa[i][p] = *reinterpret_cast<const bf8*>(&As[buf][p][(wm + i * 16 + lr) * LDK + lq * 8]);
        fb[i][p] = *reinterpret_cast<const bf8*>(&Bs[buf][p][(wn + i * 16 + lr) * LDK + lq * 8]);
      }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        // small terms first so the largest product is added last
        if (NP == 6) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][2], fb[j][0], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][1], fb[j][1], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][0], fb[j][2], acc[i][j], 0, 0, 0);
        }
        if (NP >= 3) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][1], fb[j][0], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][0], fb[j][1], acc[i][j], 0, 0, 0);
        }
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][0], fb[j][0], acc[i][j], 0, 0, 0);
      }
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // transposed output tile in LDS: Ds[n_local][m_local]; lane holds D[4*lq+v][lr]
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ml = wm + i * 16 + lq * 4 + v, nl = wn + j * 16 + lr;
        float val = acc[i][j][v];
        if (MODE == X_SPECTRAL) {  // D[m=l][n=k] = Y[k][l]
          const int r = m0 + ml, c = n0 + nl;
          const int dm = r > c ? r - c : c - r;
          const double lam = 2. * (sq[dm] + sq[r + c]);
          val = (r == 0 && c == 0) ? 0.f : float(double(val) / lam);
        }
        Ds[nl * LDD + ml] = val;
      }
  __syncthreads();
  // each thread writes 16 consecutive m of one transposed row
  const int nl = t >> 2, mc = (t & 3) * 16;
  const long orow = long(n0 + nl) * n + m0 + mc;
  double nrm[1] = {0.};
  if (MODE == X_PLAIN || MODE == X_SPECTRAL) {
    u16 h[3][16];
#pragma unroll
    for (int e = 0; e < 16; ++e) split3(Ds[nl * LDD + mc + e], h[0][e], h[1][e], h[2][e]);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      uint4* dst = reinterpret_cast<uint4*>(out16 + p * ps + orow);
      dst[0] = *reinterpret_cast<const uint4*>(&h[p][0]);
      dst[1] = *reinterpret_cast<const uint4*>(&h[p][8]);
    }
  } else {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      double v = double(Ds[nl * LDD + mc + e]);
      if (MODE == X_FINAL_ADD) v = add64[orow + e] + v;
      out64[orow + e] = v;
      nrm[0] += v * v;
    }
  }
  if ((MODE == X_FINAL || MODE == X_FINAL_ADD) && st && counter) {
    if (!publish_partials<1>(nrm, partials, counter)) return;
    fold_partials<1>(nrm, partials, counter, block_count());
    if (threadIdx.x == 0) {
      const double dn = sqrt(nrm[0]);
      st->dnorm = dn;
      if (dn <= 2 * DBL_MIN) {
        st->done = 1;
        st->exit_reason = 2;
      }
    }
  }
}

// FP64 face field -> 3 BF16 planes; element 0 gets r[0] - S (rank-1 pin correction)
__global__ __launch_bounds__(NT) void k_split_rhs(const double* __restrict__ r, u16* __restrict__ out, long F, long ps,
                                                 const SolveState* st) {
  if (st && st->done) return;
  const double S = st ? st->sum_r : 0.;
  for (long i = long(blockIdx.x) * NT + threadIdx.x; i < F; i += long(gridDim.x) * NT) {
    double v = r[i];
    if (i == 0) v -= S;
    u16 h0, h1, h2;
    split3d(v, h0, h1, h2);
    out[i] = h0;
    out[ps + i] = h1;
    out[2 * ps + i] = h2;
  }
}

// s = r' - M0 x  (FP64), written as 3 BF16 planes; r' = r - S e0
__global__ __launch_bounds__(NT) void k_refine_residual(const double* __restrict__ r, const double* __restrict__ x,
                                                       u16* __restrict__ out, int n, long ps, const SolveState* st) {
  if (st && st->done) return;
  const double S = st ? st->sum_r : 0.;
  const long F = long(n) * n;
  for (long f = long(blockIdx.x) * NT + threadIdx.x; f < F; f += long(gridDim.x) * NT) {
    const int i = int(f / n), j = int(f - (f / n) * n);
    const int i1 = i == 0 ? i : i - 1, j1 = j == 0 ? j : j - 1;
    const int i2 = i == n - 1 ? i : i + 1, j2 = j == n - 1 ? j : j + 1;
    const double nb = (x[long(i1) * n + j1] + x[long(i1) * n + j2]) + (x[long(i2) * n + j1] + x[long(i2) * n + j2]);
    double rv = r[f];
    if (f == 0) rv -= S;
    const double s = rv - (2.0 * x[f] - 0.5 * nb);
    u16 h0, h1, h2;
    split3d(s, h0, h1, h2);
    out[f] = h0;
    out[ps + f] = h1;
    out[2 * ps + f] = h2;
  }
}

template <int MODE, int NP>
void mgemm(const u16* A, const u16* B, int n, long ps, u16* out16, double* out64, const double* add64,
           const double* sq, SolveState* st, double* partials, unsigned* counter, hipStream_t s) {
  const dim3 grid(n / NB, n / MB);
  hipLaunchKernelGGL((k_gemm_bf16x3<MODE, NP>), grid, dim3(NT), 0, s, A, B, n, ps, out16, out64, add64, sq, st,
                     partials, counter);
  OTM_HIP(hipGetLastError());
}

template <int NP>
void mixed_transform(const MixedPlan& mp, const DctPlan& p, const Fields& f, const u16* rhs, double* out,
                     const double* add, bool final_add, hipStream_t s) {
  const int n = p.n;
  const long ps = long(n) * n;
  SolveState* st = f.state;
  u16* t0 = mp.work;
  u16* t1 = mp.work + 3 * ps;
  mgemm<X_PLAIN, NP>(rhs, mp.C16, n, ps, t0, nullptr, nullptr, p.sq, st, nullptr, nullptr, s);
  mgemm<X_SPECTRAL, NP>(t0, mp.C16, n, ps, t1, nullptr, nullptr, p.sq, st, nullptr, nullptr, s);
  mgemm<X_PLAIN, NP>(t1, mp.Ct16, n, ps, t0, nullptr, nullptr, p.sq, st, nullptr, nullptr, s);
  if (final_add)
    mgemm<X_FINAL_ADD, NP>(t0, mp.Ct16, n, ps, nullptr, out, add, p.sq, st, f.partials, f.counters + CNT_DNORM, s);
  else
    mgemm<X_FINAL, NP>(t0, mp.Ct16, n, ps, nullptr, out, nullptr, p.sq, st, nullptr, nullptr, s);
}

}  // namespace

bool mixed_supported(int n) { return n % 64 == 0; }

void mixed_plan_create(MixedPlan& mp, const DctPlan& p) {
  const int n = p.n;
  mp.n = n;
  if (!mixed_supported(n)) return;
  const size_t nn = size_t(n) * n;
  // DCT planes from the FP64 matrix (host copy), split in FP64
  std::vector<double> C(nn);
  OTM_HIP(hipMemcpy(C.data(), p.C, nn * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<u16> c16(3 * nn), ct16(3 * nn);
  auto bits = [](float x) {
    __bf16 h = (__bf16)x;
    return __builtin_bit_cast(u16, h);
  };
  auto val = [](u16 b) { return (double)(float)__builtin_bit_cast(__bf16, b); };
  for (size_t k = 0; k < size_t(n); ++k)
    for (size_t i = 0; i < size_t(n); ++i) {
      const double x = C[k * n + i];
      const u16 h0 = bits((float)x);
      const double r1 = x - val(h0);
      const u16 h1 = bits((float)r1);
      const double r2 = r1 - val(h1);
      const u16 h2 = bits((float)r2);
      const u16 h[3] = {h0, h1, h2};
      for (int q = 0; q < 3; ++q) {
        c16[q * nn + k * n + i] = h[q];   // C  [k][i]
        ct16[q * nn + i * n + k] = h[q];  // C^T[i][k]
      }
    }
  OTM_HIP(hipMalloc(&mp.C16, 3 * nn * sizeof(u16)));
  OTM_HIP(hipMalloc(&mp.Ct16, 3 * nn * sizeof(u16)));
  OTM_HIP(hipMalloc(&mp.work, 6 * nn * sizeof(u16)));
  OTM_HIP(hipMalloc(&mp.rhs, 3 * nn * sizeof(u16)));
  OTM_HIP(hipMalloc(&mp.x0, nn * sizeof(double)));
  OTM_HIP(hipMemcpy(mp.C16, c16.data(), 3 * nn * sizeof(u16), hipMemcpyHostToDevice));
  OTM_HIP(hipMemcpy(mp.Ct16, ct16.data(), 3 * nn * sizeof(u16), hipMemcpyHostToDevice));
}

void mixed_plan_destroy(MixedPlan& mp) {
  if (mp.C16) (void)hipFree(mp.C16);
  if (mp.Ct16) (void)hipFree(mp.Ct16);
  if (mp.work) (void)hipFree(mp.work);
  if (mp.rhs) (void)hipFree(mp.rhs);
  if (mp.x0) (void)hipFree(mp.x0);
  mp = MixedPlan();
}

// d_hat = M^+ (r - S e0): mixed transform, then `refine` FP64 refinement steps.
void launch_mixed_solve(const MixedPlan& mp, const DctPlan& p, const Fields& f, const double* r, double* out,
                        int refine, hipStream_t s) {
  const int n = p.n;
  const long F = long(n) * n, ps = F;
  const unsigned blocks = unsigned(std::min<long>((F + NT - 1) / NT, 8192));
  hipLaunchKernelGGL(k_split_rhs, dim3(blocks), dim3(NT), 0, s, r, mp.rhs, F, ps, (const SolveState*)f.state);
  OTM_HIP(hipGetLastError());
  if (refine <= 0) {
    // single transform: the final stage also publishes |d_hat|
    const long ps2 = F;
    SolveState* st = f.state;
    u16* t0 = mp.work;
    u16* t1 = mp.work + 3 * ps2;
    mgemm<X_PLAIN, 6>(mp.rhs, mp.C16, n, ps2, t0, nullptr, nullptr, p.sq, st, nullptr, nullptr, s);
    mgemm<X_SPECTRAL, 6>(t0, mp.C16, n, ps2, t1, nullptr, nullptr, p.sq, st, nullptr, nullptr, s);
    mgemm<X_PLAIN, 6>(t1, mp.Ct16, n, ps2, t0, nullptr, nullptr, p.sq, st, nullptr, nullptr, s);
    hipLaunchKernelGGL((k_gemm_bf16x3<X_FINAL, 6>), dim3(n / NB, n / MB), dim3(NT), 0, s, (const u16*)t0,
                       (const u16*)mp.Ct16, n, ps2, (u16*)nullptr, out, (const double*)nullptr, (const double*)p.sq,
                       st, f.partials, f.counters + CNT_DNORM);
    OTM_HIP(hipGetLastError());
    return;
  }
  mixed_transform<6>(mp, p, f, mp.rhs, mp.x0, nullptr, false, s);
  for (int k = 0; k < refine; ++k) {
    hipLaunchKernelGGL(k_refine_residual, dim3(blocks), dim3(NT), 0, s, r, (const double*)mp.x0, mp.rhs, n, ps,
                       (const SolveState*)f.state);
    OTM_HIP(hipGetLastError());
    const bool last = (k == refine - 1);
    if (last)
      mixed_transform<6>(mp, p, f, mp.rhs, out, mp.x0, true, s);
    else
      mixed_transform<6>(mp, p, f, mp.rhs, mp.x0, mp.x0, true, s);
  }
}

}  // namespace otm
