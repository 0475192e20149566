// Exact Laplacian solve of the reference's pinned pseudo-Laplacian by a 2-D
// DCT-II computed as dense GEMMs on FP64 MFMA (v_mfma_f64_16x16x4f64).
//
// The reference assembles M = -L with weight -1/2 at the four clamped
// diagonal neighbours and 2 on the diagonal (otlib/otsolver_2dgrid.cpp:284-311),
// pins M(0,0) (:321), factorises once and solves d = M^{-1} r followed by mean
// removal (:175-179).  Clamped indices make M = 2I - 1/2 S(x)S with S the
// 1-D Neumann operator 2I - T, so the orthonormal DCT-II C (C[k][i] =
// s_k cos(pi k (i+1/2)/n)) diagonalises it:
//   lambda_kl = 2 (1 - cos(pi k/n) cos(pi l/n))
//             = 2 (sin^2(pi (k-l)/2n) + sin^2(pi (k+l)/2n))   (no cancellation)
// with the single null mode (0,0).  Pinning then projecting onto mean zero
// equals the pseudo-inverse applied to r - (sum r) e_0 for every r and every
// pin weight, so
//   d_hat = C^T [ (C R' C^T) / lambda ] C,   R' = R - (sum r) e_0 e_0^T,
// four n x n x n GEMMs:  T1 = C R',  Y = (T1 C^T)/lambda,  T3 = C^T Y,  D = T3 C.
// (The mean of d_hat is zero up to rounding and no observable of the solver
// depends on it: every use of psi or d goes through vertex gradients.)

#include <cfloat>
#include <cmath>
#include <vector>

#include "otm_internal.h"
#include "reduce.h"

namespace otm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;

enum GemmMode { G_PLAIN = 0, G_RHS = 1, G_SPECTRAL = 2, G_FINAL = 3 };

// D = A * B, all n x n row-major.  Workgroup tile 64x64, four waves each
// owning a 32x32 quadrant as 2x2 MFMA 16x16 tiles; K staged through LDS in
// steps of 16 with register prefetch and double-buffered LDS.
//   G_RHS:      B(0,0) -= state->sum_r while loading (R' = R - S e0 e0^T)
//   G_SPECTRAL: D(k,l) /= lambda_kl, D(0,0) = 0
//   G_FINAL:    |D|^2 workgroup partials (the scalar kernel that follows stops
//               the solve when |d_hat| <= 2*DBL_MIN, otsolver_2dgrid.cpp:182)
template <int MODE, bool EDGE>
__global__ __launch_bounds__(NT) void k_gemm_f64(const double* __restrict__ A, const double* __restrict__ B,
                                                double* __restrict__ D, int n, SolveState* st,
                                                const double* __restrict__ sq, double* partials) {
  if (st && st->done) return;
  __shared__ double As[2][BM][BK + 1];
  __shared__ double Bs[2][BK][BN + 2];
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  const int ar = t >> 2, ak = (t & 3) * 4;   // A loads: row ar, k ak..ak+3
  const int bk = t >> 4, bc = (t & 15) * 4;  // B loads: k bk, cols bc..bc+3
  const double S = (MODE == G_RHS && st) ? st->sum_r : 0.;
  double ra[4], rb[4];

  auto gload = [&](int k0) {
    const int gm = m0 + ar;
    const int gk = k0 + bk;
    if (!EDGE) {
      const double2* pa = reinterpret_cast<const double2*>(A + long(gm) * n + k0 + ak);
      const double2 x0 = pa[0], x1 = pa[1];
      ra[0] = x0.x; ra[1] = x0.y; ra[2] = x1.x; ra[3] = x1.y;
      const double2* pb = reinterpret_cast<const double2*>(B + long(gk) * n + n0 + bc);
      const double2 y0 = pb[0], y1 = pb[1];
      rb[0] = y0.x; rb[1] = y0.y; rb[2] = y1.x; rb[3] = y1.y;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ka = k0 + ak + q;
        ra[q] = (gm < n && ka < n) ? A[long(gm) * n + ka] : 0.;
        const int gc = n0 + bc + q;
        rb[q] = (gk < n && gc < n) ? B[long(gk) * n + gc] : 0.;
      }
    }
    if (MODE == G_RHS && gk == 0 && n0 + bc == 0) rb[0] -= S;
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[buf][ar][ak + q] = ra[q];
      Bs[buf][bk][bc + q] = rb[q];
    }
  };

  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0., 0., 0., 0.};

  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < n; k0 += BK) {
    const bool more = k0 + BK < n;
    if (more) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const double a0 = As[buf][wm + lr][kk + lk];
      const double a1 = As[buf][wm + 16 + lr][kk + lk];
      const double b0 = Bs[buf][kk + lk][wn + lr];
      const double b1 = Bs[buf][kk + lk][wn + 16 + lr];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // epilogue: for v_mfma_f64_16x16x4f64 lane l holds D[l/16 + 4v][l%16] of each 16x16 tile
  double nrm[1] = {0.};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + wm + i * 16 + lk + 4 * v;
        const int col = n0 + wn + j * 16 + lr;
        if (EDGE && (row >= n || col >= n)) continue;
        double val = acc[i][j][v];
        if (MODE == G_SPECTRAL) {
          const int dm = row > col ? row - col : col - row;
          const double lam = 2. * (sq[dm] + sq[row + col]);
          val = (row == 0 && col == 0) ? 0. : val / lam;
        }
        if (MODE == G_FINAL) nrm[0] += val * val;
        D[long(row) * n + col] = val;
      }
  if (MODE == G_FINAL && partials) write_partials<1>(nrm, partials);
}

template <int MODE>
void gemm(const double* A, const double* B, double* D, int n, SolveState* st, const double* sq, double* partials,
          hipStream_t s) {
  const dim3 grid((n + BN - 1) / BN, (n + BM - 1) / BM);
  if (n % 64 == 0)
    hipLaunchKernelGGL((k_gemm_f64<MODE, false>), grid, dim3(NT), 0, s, A, B, D, n, st, sq, partials);
  else
    hipLaunchKernelGGL((k_gemm_f64<MODE, true>), grid, dim3(NT), 0, s, A, B, D, n, st, sq, partials);
  OTM_HIP(hipGetLastError());
}

}  // namespace

void dct_plan_create(DctPlan& p, int n) {
  p.n = n;
  const size_t nn = size_t(n) * n;
  std::vector<double> C(nn), Ct(nn), sq(2 * size_t(n) + 1);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < n; ++k) {
    const long double sk = k == 0 ? std::sqrt(1.0L / n) : std::sqrt(2.0L / n);
    for (int i = 0; i < n; ++i) {
      // angle pi*k*(2i+1)/(2n) reduced exactly modulo 2*pi: m = k*(2i+1) mod 4n
      const long m = (long(k) * (2 * i + 1)) % (4L * n);
      const double c = double(sk * std::cos(pi * (long double)m / (2.0L * n)));
      C[size_t(k) * n + i] = c;
      Ct[size_t(i) * n + k] = c;
    }
  }
  for (int m = 0; m <= 2 * n; ++m) {
    const long double sv = std::sin(pi * (long double)m / (2.0L * n));
    sq[m] = double(sv * sv);
  }
  OTM_HIP(hipMalloc(&p.C, nn * sizeof(double)));
  OTM_HIP(hipMalloc(&p.Ct, nn * sizeof(double)));
  OTM_HIP(hipMalloc(&p.sq, sq.size() * sizeof(double)));
  OTM_HIP(hipMemcpy(p.C, C.data(), nn * sizeof(double), hipMemcpyHostToDevice));
  OTM_HIP(hipMemcpy(p.Ct, Ct.data(), nn * sizeof(double), hipMemcpyHostToDevice));
  OTM_HIP(hipMemcpy(p.sq, sq.data(), sq.size() * sizeof(double), hipMemcpyHostToDevice));
}

void dct_plan_destroy(DctPlan& p) {
  if (p.C) (void)hipFree(p.C);
  if (p.Ct) (void)hipFree(p.Ct);
  if (p.sq) (void)hipFree(p.sq);
  p = DctPlan();
}

// one plain stage (T3 = C^T Y) without gating, for kernel timing
void launch_dct_gemm_stage(const DctPlan& p, const Fields& f, hipStream_t s) {
  gemm<G_PLAIN>(p.Ct, f.tmp1, f.tmp0, p.n, nullptr, p.sq, nullptr, s);
}

// d_hat = C^T [(C R' C^T) / lambda] C.  Every GEMM is skipped once state->done
// is set; the last stage leaves the |d_hat|^2 partials in SITE_DNORM.
void launch_dct_solve(const DctPlan& p, const Fields& f, const double* r, double* out, hipStream_t s) {
  const int n = p.n;
  SolveState* st = f.state;
  gemm<G_RHS>(p.C, r, f.tmp0, n, st, p.sq, nullptr, s);
  gemm<G_SPECTRAL>(f.tmp0, p.Ct, f.tmp1, n, st, p.sq, nullptr, s);
  gemm<G_PLAIN>(p.Ct, f.tmp1, f.tmp0, n, st, p.sq, nullptr, s);
  gemm<G_FINAL>(f.tmp0, p.C, out, n, st, p.sq, f.site(SITE_DNORM), s);
}

}  // namespace otm
