// Mixed-precision Laplacian solve: the same 2-D DCT-II diagonalisation as the
// FP64 path, with every GEMM run as a 3-term BF16 split on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), followed by FP64 iterative
// refinement against the exact operator.
//
// Operands and intermediates live in HBM as FP32 (4 bytes per element).  A
// workgroup splits each FP32 value exactly into three BF16 terms
// x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1))
// while staging it into LDS, and a product keeps the six terms a_i b_j with
// i + j <= 2 (NP = 6; NP = 3 keeps i + j <= 1).
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
// Kernel shape: 64x64 output tile per workgroup (four waves, 32x32 each as
// 2x2 MFMA tiles), K in steps of 64 with a 3-deep register prefetch ring and
// double-buffered LDS, and an XCD-aware tile order: workgroups are dealt to
// the 8 XCDs round-robin, so consecutive ids that land on one XCD get a
// compact 4 x 8 block of tiles whose operand panels fit that XCD's L2.
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

constexpr int TM = 64, TN = 64, KB = 64;  // workgroup tile and K step
constexpr int DEPTH = 3;                  // register prefetch ring
constexpr int LDK = KB + 8;               // LDS row length (bf16) of operand tiles: 144 B, conflict-free reads
constexpr int LDD = TM + 4;               // LDS row length (f32) of the transposed output tile
constexpr int PLANE = TM * LDK;           // u16 per plane of one operand tile

enum MixedMode { X_PLAIN = 0, X_SPECTRAL = 1, X_FINAL = 2, X_FINAL_ADD = 3 };

__device__ inline u16 bf16_bits(float x) {
  __bf16 h = (__bf16)x;
  return __builtin_bit_cast(u16, h);
}
__device__ inline float bf16_float(u16 b) { return (float)__builtin_bit_cast(__bf16, b); }

// exact 3-term split of an FP32 value
__device__ inline void split3(float x, u16& h0, u16& h1, u16& h2) {
  h0 = bf16_bits(x);
  const float r1 = x - bf16_float(h0);
  h1 = bf16_bits(r1);
  const float r2 = r1 - bf16_float(h1);
  h2 = bf16_bits(r2);
}

// tile coordinates for workgroup id b of an nt x nt grid (XCD-aware when nt % 4 == 0)
__device__ inline void tile_of(int b, int nt, int& tm, int& tn) {
  if (nt % 4 == 0 && nt >= 8) {
    const int per = nt * nt / 8;          // tiles per XCD
    const int xcd = b % 8, local = b / 8;
    const int cols = nt / 2, rows = per / cols;  // rows x cols block of tiles per XCD
    const int bx = xcd % 2, by = xcd / 2;
    tm = by * rows + local / cols;
    tn = bx * cols + local % cols;
  } else {
    tm = b / nt;
    tn = b % nt;
  }
}

// D = A * B^T, A and B n x n FP32 row-major (k contiguous).
template <int MODE, int NP>
__global__ __launch_bounds__(NT) void k_gemm_x3(const float* __restrict__ A, const float* __restrict__ B, int n,
                                               float* __restrict__ out32, double* __restrict__ out64,
                                               const double* __restrict__ add64, const double* __restrict__ sq,
                                               const SolveState* st, double* partials) {
  if (st && st->done) return;
  constexpr int NPL = (NP == 1) ? 1 : (NP == 3 ? 2 : 3);  // planes used
  // [buf][operand][plane][row * LDK + k]; the output tile reuses the same storage
  __shared__ __attribute__((aligned(16))) u16 lds[2 * 2 * 3 * PLANE];
  const int nt = n / TM;
  int tm, tn;
  tile_of(blockIdx.x, nt, tm, tn);
  const int m0 = tm * TM, n0 = tn * TN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lr = lane & 15, lq = lane >> 4;
  const int grow = t >> 2, gk = (t & 3) * 16;  // staging: row, 16 consecutive k
  const float* pa = A + long(m0 + grow) * n + gk;
  const float* pb = B + long(n0 + grow) * n + gk;
  const int nsteps = n / KB;

  float4 ra[DEPTH][4], rb[DEPTH][4];
  auto gload = [&](int slot, int step) {
    const float4* qa = reinterpret_cast<const float4*>(pa + step * KB);
    const float4* qb = reinterpret_cast<const float4*>(pb + step * KB);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra[slot][q] = qa[q];
      rb[slot][q] = qb[q];
    }
  };
  auto sstore = [&](int slot, int buf) {
    u16* la = lds + (buf * 2 + 0) * 3 * PLANE + grow * LDK + gk;
    u16* lb = lds + (buf * 2 + 1) * 3 * PLANE + grow * LDK + gk;
    u16 ha[3][16], hb[3][16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float xa[4] = {ra[slot][q].x, ra[slot][q].y, ra[slot][q].z, ra[slot][q].w};
      const float xb[4] = {rb[slot][q].x, rb[slot][q].y, rb[slot][q].z, rb[slot][q].w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        split3(xa[e], ha[0][4 * q + e], ha[1][4 * q + e], ha[2][4 * q + e]);
        split3(xb[e], hb[0][4 * q + e], hb[1][4 * q + e], hb[2][4 * q + e]);
      }
    }
#pragma unroll
    for (int p = 0; p < NPL; ++p) {
      uint4 va0, va1, vb0, vb1;
      va0.x = ha[p][0] | (unsigned(ha[p][1]) << 16); va0.y = ha[p][2] | (unsigned(ha[p][3]) << 16);
      va0.z = ha[p][4] | (unsigned(ha[p][5]) << 16); va0.w = ha[p][6] | (unsigned(ha[p][7]) << 16);
      va1.x = ha[p][8] | (unsigned(ha[p][9]) << 16); va1.y = ha[p][10] | (unsigned(ha[p][11]) << 16);
      va1.z = ha[p][12] | (unsigned(ha[p][13]) << 16); va1.w = ha[p][14] | (unsigned(ha[p][15]) << 16);
      vb0.x = hb[p][0] | (unsigned(hb[p][1]) << 16); vb0.y = hb[p][2] | (unsigned(hb[p][3]) << 16);
      vb0.z = hb[p][4] | (unsigned(hb[p][5]) << 16); vb0.w = hb[p][6] | (unsigned(hb[p][7]) << 16);
      vb1.x = hb[p][8] | (unsigned(hb[p][9]) << 16); vb1.y = hb[p][10] | (unsigned(hb[p][11]) << 16);
      vb1.z = hb[p][12] | (unsigned(hb[p][13]) << 16); vb1.w = hb[p][14] | (unsigned(hb[p][15]) << 16);
      *reinterpret_cast<uint4*>(la + p * PLANE) = va0;
      *reinterpret_cast<uint4*>(la + p * PLANE + 8) = va1;
      *reinterpret_cast<uint4*>(lb + p * PLANE) = vb0;
      *reinterpret_cast<uint4*>(lb + p * PLANE + 8) = vb1;
    }
  };

  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  auto compute = [&](int buf) {
    const u16* la = lds + (buf * 2 + 0) * 3 * PLANE;
    const u16* lb = lds + (buf * 2 + 1) * 3 * PLANE;
#pragma unroll
    for (int kk = 0; kk < KB; kk += 32) {
      bf8 fa[2][3], fb[2][3];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int p = 0; p < NPL; ++p) {
          fa[i][p] = *reinterpret_cast<const bf8*>(la + p * PLANE + (wm + i * 16 + lr) * LDK + kk + lq * 8);
          fb[i][p] = *reinterpret_cast<const bf8*>(lb + p * PLANE + (wn + i * 16 + lr) * LDK + kk + lq * 8);
        }
      // product-major order: consecutive MFMAs update different accumulators;
      // small terms first, the leading product last
      constexpr int PA[6] = {2, 1, 0, 1, 0, 0};
      constexpr int PB[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
      for (int q = 6 - NP; q < 6; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][PA[q]], fb[j][PB[q]], acc[i][j], 0, 0, 0);
    }
  };

#pragma unroll
  for (int s = 0; s < DEPTH; ++s)
    if (s < nsteps) gload(s, s);
  for (int s0 = 0; s0 < nsteps; s0 += DEPTH) {
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
      const int s = s0 + u;
      if (s < nsteps) {
        const int buf = s & 1;
        sstore(u, buf);
        if (s + DEPTH < nsteps) gload(u, s + DEPTH);
        __syncthreads();
        compute(buf);
      }
    }
  }
  __syncthreads();

  // transposed output tile in LDS: Ds[n_local][m_local]; lane holds D[4*lq+v][lr]
  float* Ds = reinterpret_cast<float*>(lds);
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
  if (MODE == X_PLAIN || MODE == X_SPECTRAL) {
    float4* dst = reinterpret_cast<float4*>(out32 + orow);
    const float4* src = reinterpret_cast<const float4*>(Ds + nl * LDD + mc);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = src[q];
  } else {
    double nrm[1] = {0.};
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      double v = double(Ds[nl * LDD + mc + e]);
      if (MODE == X_FINAL_ADD) v = add64[orow + e] + v;
      out64[orow + e] = v;
      nrm[0] += v * v;
    }
    if (partials) write_partials<1>(nrm, partials);
  }
}

// FP64 face field -> FP32 right-hand side; element 0 gets r[0] - S (rank-1 pin correction)
__global__ __launch_bounds__(NT) void k_split_rhs(const double* __restrict__ r, float* __restrict__ out, long F,
                                                 const SolveState* st) {
  if (st && st->done) return;
  const double S = st ? st->sum_r : 0.;
  for (long i = long(blockIdx.x) * NT + threadIdx.x; i < F; i += long(gridDim.x) * NT) {
    double v = r[i];
    if (i == 0) v -= S;
    out[i] = float(v);
  }
}

// s = r' - M0 x  (FP64), stored as FP32; r' = r - S e0
__global__ __launch_bounds__(NT) void k_refine_residual(const double* __restrict__ r, const double* __restrict__ x,
                                                       float* __restrict__ out, int n, const SolveState* st) {
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
    out[f] = float(rv - (2.0 * x[f] - 0.5 * nb));
  }
}

template <int MODE, int NP>
void mgemm(const float* A, const float* B, int n, float* out32, double* out64, const double* add64, const double* sq,
           const SolveState* st, double* partials, hipStream_t s) {
  const int nt = n / TM;
  hipLaunchKernelGGL((k_gemm_x3<MODE, NP>), dim3(nt * nt), dim3(NT), 0, s, A, B, n, out32, out64, add64, sq, st,
                     partials);
  OTM_HIP(hipGetLastError());
}

// out = (add +) transform(rhs); publish: the final stage leaves the |out|^2
// partials in SITE_DNORM for the early-stop test (otsolver_2dgrid.cpp:182)
template <int NP>
void mixed_transform(const MixedPlan& mp, const DctPlan& p, const Fields& f, const float* rhs, double* out,
                     const double* add, bool publish, hipStream_t s) {
  const int n = p.n;
  const long nn = long(n) * n;
  const SolveState* st = f.state;
  float* t0 = mp.work;
  float* t1 = mp.work + nn;
  double* part = publish ? f.site(SITE_DNORM) : nullptr;
  mgemm<X_PLAIN, NP>(rhs, mp.C32, n, t0, nullptr, nullptr, p.sq, st, nullptr, s);
  mgemm<X_SPECTRAL, NP>(t0, mp.C32, n, t1, nullptr, nullptr, p.sq, st, nullptr, s);
  mgemm<X_PLAIN, NP>(t1, mp.Ct32, n, t0, nullptr, nullptr, p.sq, st, nullptr, s);
  if (add)
    mgemm<X_FINAL_ADD, NP>(t0, mp.Ct32, n, nullptr, out, add, p.sq, st, part, s);
  else
    mgemm<X_FINAL, NP>(t0, mp.Ct32, n, nullptr, out, nullptr, p.sq, st, part, s);
}

}  // namespace

bool mixed_supported(int n) { return n % 64 == 0; }

void mixed_plan_create(MixedPlan& mp, const DctPlan& p) {
  const int n = p.n;
  mp.n = n;
  if (!mixed_supported(n)) return;
  const size_t nn = size_t(n) * n;
  std::vector<double> C(nn), Ct(nn);
  OTM_HIP(hipMemcpy(C.data(), p.C, nn * sizeof(double), hipMemcpyDeviceToHost));
  OTM_HIP(hipMemcpy(Ct.data(), p.Ct, nn * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<float> c32(nn), ct32(nn);
  for (size_t k = 0; k < nn; ++k) {
    c32[k] = float(C[k]);
    ct32[k] = float(Ct[k]);
  }
  OTM_HIP(hipMalloc(&mp.C32, nn * sizeof(float)));
  OTM_HIP(hipMalloc(&mp.Ct32, nn * sizeof(float)));
  OTM_HIP(hipMalloc(&mp.work, 2 * nn * sizeof(float)));
  OTM_HIP(hipMalloc(&mp.rhs, nn * sizeof(float)));
  OTM_HIP(hipMalloc(&mp.x0, nn * sizeof(double)));
  OTM_HIP(hipMemcpy(mp.C32, c32.data(), nn * sizeof(float), hipMemcpyHostToDevice));
  OTM_HIP(hipMemcpy(mp.Ct32, ct32.data(), nn * sizeof(float), hipMemcpyHostToDevice));
}

void mixed_plan_destroy(MixedPlan& mp) {
  if (mp.C32) (void)hipFree(mp.C32);
  if (mp.Ct32) (void)hipFree(mp.Ct32);
  if (mp.work) (void)hipFree(mp.work);
  if (mp.rhs) (void)hipFree(mp.rhs);
  if (mp.x0) (void)hipFree(mp.x0);
  mp = MixedPlan();
}

// one ungated 6-product stage (stage 3 shape), for kernel timing
void launch_mixed_gemm_stage(const MixedPlan& mp, const DctPlan& p, hipStream_t s) {
  const long nn = long(p.n) * p.n;
  mgemm<X_PLAIN, 6>(mp.work + nn, mp.Ct32, p.n, mp.work, nullptr, nullptr, p.sq, nullptr, nullptr, s);
}

// d_hat = M^+ (r - S e0): mixed transform, then `refine` FP64 refinement steps
// (the correction transforms keep NP_CORR products).
void launch_mixed_solve(const MixedPlan& mp, const DctPlan& p, const Fields& f, const double* r, double* out,
                        int refine, hipStream_t s) {
  const int n = p.n;
  const long F = long(n) * n;
  const unsigned blocks = unsigned(std::min<long>((F + NT - 1) / NT, 8192));
  hipLaunchKernelGGL(k_split_rhs, dim3(blocks), dim3(NT), 0, s, r, mp.rhs, F, (const SolveState*)f.state);
  OTM_HIP(hipGetLastError());
  if (refine <= 0) {
    mixed_transform<6>(mp, p, f, mp.rhs, out, nullptr, true, s);
    return;
  }
  mixed_transform<6>(mp, p, f, mp.rhs, mp.x0, nullptr, false, s);
  for (int k = 0; k < refine; ++k) {
    hipLaunchKernelGGL(k_refine_residual, dim3(blocks), dim3(NT), 0, s, r, (const double*)mp.x0, mp.rhs, n,
                       (const SolveState*)f.state);
    OTM_HIP(hipGetLastError());
    const bool last = (k == refine - 1);
    // all six products: with the three leading ones only, one refinement step
    // leaves ~1e-9 instead of ~2e-10 relative error at n = 1024
    mixed_transform<6>(mp, p, f, mp.rhs, last ? out : mp.x0, mp.x0, last, s);
  }
}

}  // namespace otm
