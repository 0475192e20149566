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

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "otm_internal.h"
#include "reduce.h"

namespace otm {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double dv2 __attribute__((ext_vector_type(2)));

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

// ---------------------------------------------------------------------------
// Even/odd folded transform (n % 64 == 0, h = n/2).  Row k of C is symmetric
// (k even) or antisymmetric (k odd) under i -> n-1-i, so with P = [E; O]
// (E, O = even / odd rows of C restricted to i < h) the four stages become
//   F1  T  = P  . fold_rows(R')           rows of T in (even k, odd k) order
//   F2  Y  = fold_cols(T) . P^T / lambda  cols of Y in (even l, odd l) order
//   F3  T3 = P^T . Y     rows i and n-1-i from the two K halves
//   F4  D  = T3 . P      cols j and n-1-j from the two K halves
// where fold(x)[i] = x[i] + s x[n-1-i] with s = +1 for the even half of the
// output and -1 for the odd half.  Each stage is n^3 flops instead of 2 n^3.
// Workgroup tile FM x FN (four waves, 2 x 2), K step KB, double-buffered LDS.
// Every sum is taken in two halves of K, each in increasing k, added (F1, F2,
// F3 rows i / F4 columns j) or subtracted (F3 rows n-1-i / F4 columns n-1-j)
// at the end: F3 and F4 split at k = h (the fold needs both halves), F1 and
// F2 at k = h/2 when n % 128 == 0 (else in one piece).  With G = 1 one group
// of four waves runs both halves in turn; with G = 2 (K halves split between
// two groups of four waves, 512 threads) each group runs one half and the
// second group's sums cross over LDS, so a 64 x 64 tile keeps eight waves
// busy.  The tile shape is picked per stage by dct_tune (every shape sums in
// the same order, so every shape gives bit-identical results).
enum FoldStage { F1 = 0, F2 = 1, F3 = 2, F4 = 3 };

template <int ST, int FM, int FN, int KB, int G>
__global__ __launch_bounds__(G * NT) void k_gemm_fold(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                     int count, bool gate, bool rhs) {
  const int p = count == 1 ? 0 : job_of(starts, count, blockIdx.x);
  const int lt = count == 1 ? int(blockIdx.x) : blockIdx.x - starts[p];
  const int nb = count == 1 ? int(gridDim.x) : starts[p + 1] - starts[p];
  const Job& J = jobs[p];
  const auto st = gp((const SolveState*)J.f.state);
  if (gate && st->done) return;
  const int n = J.f.n, h = n >> 1;
  // stage operands:  F1: P (n x h) . fold_rows(r), K = h     F2: fold_cols(T) . Pt (h x n), K = h
  //                  F3: Pt rows < h . Y, K = n                F4: T3 . P cols < h, K = n
  const auto A = gp(ST == F1 ? J.P : ST == F3 ? J.Pt : (const double*)J.f.tmp0);
  const auto B = gp(ST == F1 ? (const double*)J.f.r : ST == F2 ? J.Pt : ST == F3 ? (const double*)J.f.tmp1 : J.P);
  const auto D = gp(ST == F1 ? J.f.tmp0 : ST == F2 ? J.f.tmp1 : ST == F3 ? J.f.tmp0 : J.f.dhat);
  const int lda = ST == F1 ? h : n, ldb = ST == F4 ? h : n;
  const int K = (ST == F1 || ST == F2) ? h : n;
  const int cols = ST == F4 ? h : n;
  const auto sq = gp(J.sq);
  constexpr int TI = FM / 32, TJ = FN / 32;  // each wave: FM/2 x FN/2 = TI x TJ MFMA 16x16 tiles
  constexpr int ACHUNKS = FM * KB / 4, BCHUNKS = KB * FN / 4;  // chunks of 4 consecutive values
  constexpr int ACH = (ACHUNKS + NT - 1) / NT, BCH = (BCHUNKS + NT - 1) / NT;
  constexpr bool SPLIT = ST == F3 || ST == F4;
  // LDS per group: A [2][FM][KB + 1], B [2][KB][FN + 2]; with G = 2 the
  // epilogue reuses the space for the second group's FP64 tile [FM][FN + 1]
  constexpr int A_D = 2 * FM * (KB + 1), B_D = 2 * KB * (FN + 2), G_D = A_D + B_D;
  static_assert(G == 1 || G * G_D >= FM * (FN + 1), "epilogue staging must fit");
  __shared__ double smem[G * G_D];
  const int ntn = cols / FN, ntm = (ST == F3 ? h : n) / FM;
  if (lt >= ntm * ntn) {  // padding workgroup of the job: neutral partial only
    if (ST == F4) {
      double z[1] = {0.};
      write_partials_at<1, G * NT>(z, J.f.site(SITE_DNORM), lt, nb);
    }
    return;
  }
  int tmi, tni;
  xcd_tile(lt, ntm, ntn, tmi, tni);
  const int m0 = tmi * FM, n0 = tni * FN;
  const int g = G == 2 ? int(threadIdx.x) / NT : 0;
  const int t = threadIdx.x % NT, lane = t & 63, w = t >> 6;
  double* const gs = smem + g * G_D;
  auto As = [&](int buf, int r, int c) -> double& { return gs[(buf * FM + r) * (KB + 1) + c]; };
  auto Bs = [&](int buf, int r, int c) -> double& { return gs[A_D + (buf * KB + r) * (FN + 2) + c]; };
  const int wm = (w >> 1) * (FM / 2), wn = (w & 1) * (FN / 2);
  const int lr = lane & 15, lk = lane >> 4;
  const double sgnA = (ST == F2 && n0 >= h) ? -1. : 1.;
  const double sgnB = (ST == F1 && m0 >= h) ? -1. : 1.;
  const double S = (ST == F1 && rhs) ? st->sum_r : 0.;
  // two register slots of K-step loads in flight (see the pipeline below);
  // chunk indices past the tile reload an earlier chunk (never stored) so
  // that every lane issues the same loads
  double ra[2][ACH][4], rb[2][BCH][4];

  auto gload = [&](double (&xa)[ACH][4], double (&xb)[BCH][4], int k0) {
#pragma unroll
    for (int u = 0; u < ACH; ++u) {
      const int c0 = t + u * NT, c = c0 < ACHUNKS ? c0 : c0 - ACHUNKS;
      const int ar = c / (KB / 4), ak = (c % (KB / 4)) * 4;
      const OTM_GLOBAL dv2* pa = reinterpret_cast<const OTM_GLOBAL dv2*>(A + long(m0 + ar) * lda + k0 + ak);
      const dv2 x0 = pa[0], x1 = pa[1];
      xa[u][0] = x0.x; xa[u][1] = x0.y; xa[u][2] = x1.x; xa[u][3] = x1.y;
      if (ST == F2) {  // mirrored columns n-1-(k0+ak+q), q = 0..3, read as one ascending run
        const OTM_GLOBAL dv2* pm = reinterpret_cast<const OTM_GLOBAL dv2*>(A + long(m0 + ar) * lda + (n - 4 - k0 - ak));
        const dv2 y0 = pm[0], y1 = pm[1];
        xa[u][0] += sgnA * y1.y; xa[u][1] += sgnA * y1.x; xa[u][2] += sgnA * y0.y; xa[u][3] += sgnA * y0.x;
      }
    }
#pragma unroll
    for (int u = 0; u < BCH; ++u) {
      const int c0 = t + u * NT, c = c0 < BCHUNKS ? c0 : c0 - BCHUNKS;
      const int bk = c / (FN / 4), bc = (c % (FN / 4)) * 4;
      const int gk = k0 + bk;
      const OTM_GLOBAL dv2* pb = reinterpret_cast<const OTM_GLOBAL dv2*>(B + long(gk) * ldb + n0 + bc);
      const dv2 y0 = pb[0], y1 = pb[1];
      xb[u][0] = y0.x; xb[u][1] = y0.y; xb[u][2] = y1.x; xb[u][3] = y1.y;
      if (ST == F1) {  // mirrored row n-1-gk
        const OTM_GLOBAL dv2* pm = reinterpret_cast<const OTM_GLOBAL dv2*>(B + long(n - 1 - gk) * ldb + n0 + bc);
        const dv2 z0 = pm[0], z1 = pm[1];
        xb[u][0] += sgnB * z0.x; xb[u][1] += sgnB * z0.y; xb[u][2] += sgnB * z1.x; xb[u][3] += sgnB * z1.y;
        if (gk == 0 && n0 + bc == 0) xb[u][0] -= S;  // R' = R - (sum r) e0 e0^T
      }
    }
  };
  auto sstore = [&](const double (&xa)[ACH][4], const double (&xb)[BCH][4], int buf) {
#pragma unroll
    for (int u = 0; u < ACH; ++u) {
      const int c = t + u * NT;
      if (ACHUNKS % NT == 0 || c < ACHUNKS) {
        const int ar = c / (KB / 4), ak = (c % (KB / 4)) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) As(buf, ar, ak + q) = xa[u][q];
      }
    }
#pragma unroll
    for (int u = 0; u < BCH; ++u) {
      const int c = t + u * NT;
      if (BCHUNKS % NT == 0 || c < BCHUNKS) {
        const int bk = c / (FN / 4), bc = (c % (FN / 4)) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) Bs(buf, bk, bc + q) = xb[u][q];
      }
    }
  };
  auto compute = [&](d4 (&c)[TI][TJ], int buf) {
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      double a[TI], b[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) a[i] = As(buf, wm + 16 * i + lr, kk + lk);
#pragma unroll
      for (int j = 0; j < TJ; ++j) b[j] = Bs(buf, kk + lk, wn + 16 * j + lr);
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) c[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], c[i][j], 0, 0, 0);
    }
  };

  // Software pipeline, two K-steps in flight: step s stores register slot
  // s & 1 (loaded two steps earlier) into LDS buffer s & 1, refills the slot
  // with step s + 2, syncs and multiplies; the buffer written at step s was
  // last read at step s - 2, before every wave passed the sync of step s - 1.
  // Loads past the group's last step re-read its last K block (in bounds,
  // never stored).  With G = 1 the first-half sums move to acc_lo when step
  // ksplit / KB starts (the k order of every accumulation is unchanged); with
  // G = 2 group g runs the steps of half g into acc.
  d4 acc[TI][TJ], acc_lo[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      acc[i][j] = d4{0., 0., 0., 0.};
      acc_lo[i][j] = d4{0., 0., 0., 0.};
    }
  const int ksplit = SPLIT ? h : (n % 128 == 0 ? h / 2 : K);
  const int hsteps = G == 1 ? ksplit / KB : -1;
  const int kbeg = G == 2 && g == 1 ? ksplit : 0, kend = G == 2 && g == 0 ? ksplit : K;
  const int nsteps = (kend - kbeg) / KB;
  const int klast = kend - KB;
  auto step = [&](double (&xa)[ACH][4], double (&xb)[BCH][4], int s, int buf) {
    if (G == 1 && s == hsteps) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          acc_lo[i][j] = acc[i][j];
          acc[i][j] = d4{0., 0., 0., 0.};
        }
    }
    sstore(xa, xb, buf);
    gload(xa, xb, min(kbeg + (s + 2) * KB, klast));
    __syncthreads();
    compute(acc, buf);
  };
  gload(ra[0], rb[0], kbeg);
  gload(ra[1], rb[1], min(kbeg + KB, klast));
  int s = 0;
  for (; s + 2 <= nsteps; s += 2) {
    step(ra[0], rb[0], s, 0);
    step(ra[1], rb[1], s + 1, 1);
  }
  if (s < nsteps) step(ra[0], rb[0], s, 0);

  // lane l holds D[l/16 + 4v][l%16] of each 16x16 tile.  The two halves of
  // every sum: lo (k < ksplit) and hi (k >= ksplit); with G = 2 the second
  // group hands its sums (hi) to the first over LDS
  if (G == 2) {
    __syncthreads();  // every wave is done with its operand buffers
    if (g == 1)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) smem[(wm + 16 * i + lk + 4 * v) * (FN + 1) + wn + 16 * j + lr] = acc[i][j][v];
    __syncthreads();
  }
  double nrm[1] = {0.};
  if (G == 1 || g == 0)
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int tr = wm + 16 * i + lk + 4 * v, tc = wn + 16 * j + lr;
          const int row = m0 + tr, col = n0 + tc;
          const double lo = G == 2 ? acc[i][j][v] : acc_lo[i][j][v];
          const double hi = G == 2 ? smem[tr * (FN + 1) + tc] : acc[i][j][v];
          if (ST == F1) {
            D[long(row) * n + col] = ksplit < K ? lo + hi : hi;
          } else if (ST == F2) {
            const int k = row < h ? 2 * row : 2 * (row - h) + 1;
            const int l = col < h ? 2 * col : 2 * (col - h) + 1;
            const int dm = k > l ? k - l : l - k;
            const double lam = 2. * (sq[dm] + sq[k + l]);
            D[long(row) * n + col] = (k == 0 && l == 0) ? 0. : (ksplit < K ? lo + hi : hi) / lam;
          } else {
            const double s0 = lo + hi, s1 = lo - hi;
            if (ST == F3) {
              D[long(row) * n + col] = s0;
              D[long(n - 1 - row) * n + col] = s1;
            } else {
              D[long(row) * n + col] = s0;
              D[long(row) * n + (n - 1 - col)] = s1;
              nrm[0] += s0 * s0 + s1 * s1;
            }
          }
        }
  if (ST == F4) write_partials_at<1, G * NT>(nrm, J.f.site(SITE_DNORM), lt, nb);
}

template <int ST, int CFG>
void launch_fold_cfg(const JobSet& js, bool gate, bool rhs, hipStream_t s) {
  constexpr int FM = FOLD_CFG[CFG][0], FN = FOLD_CFG[CFG][1], KB = FOLD_CFG[CFG][2], G = FOLD_CFG[CFG][3];
  const int kind = fold_kind(ST, CFG);
  if (js.total[kind] == 0) return;
  hipLaunchKernelGGL((k_gemm_fold<ST, FM, FN, KB, G>), dim3(unsigned(js.total[kind])), dim3(G * NT), 0, s,
                     (const Job*)js.jobs, (const int*)js.starts_of(kind), js.count, gate, rhs);
  OTM_HIP(hipGetLastError());
}

typedef void (*FoldLauncher)(const JobSet&, bool, bool, hipStream_t);

template <int ST, int... C>
constexpr std::array<FoldLauncher, sizeof...(C)> fold_row(std::integer_sequence<int, C...>) {
  return {&launch_fold_cfg<ST, C>...};
}

const std::array<std::array<FoldLauncher, NFOLD>, 4> FOLD_TABLE = {
    fold_row<F1>(std::make_integer_sequence<int, NFOLD>{}), fold_row<F2>(std::make_integer_sequence<int, NFOLD>{}),
    fold_row<F3>(std::make_integer_sequence<int, NFOLD>{}), fold_row<F4>(std::make_integer_sequence<int, NFOLD>{})};

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

  p.fold = n % 64 == 0;
  if (p.fold) {
    const int h = n / 2;
    std::vector<double> P(size_t(n) * h), Pt(size_t(h) * n);
    for (int r = 0; r < n; ++r) {
      const int k = r < h ? 2 * r : 2 * (r - h) + 1;
      for (int i = 0; i < h; ++i) {
        P[size_t(r) * h + i] = C[size_t(k) * n + i];
        Pt[size_t(i) * n + r] = C[size_t(k) * n + i];
      }
    }
    OTM_HIP(hipMalloc(&p.P, P.size() * sizeof(double)));
    OTM_HIP(hipMalloc(&p.Pt, Pt.size() * sizeof(double)));
    OTM_HIP(hipMemcpy(p.P, P.data(), P.size() * sizeof(double), hipMemcpyHostToDevice));
    OTM_HIP(hipMemcpy(p.Pt, Pt.data(), Pt.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  fft_plan_create(p.fft, n);
}

void dct_plan_destroy(DctPlan& p) {
  if (p.fft) (void)hipFree(p.fft);
  if (p.C) (void)hipFree(p.C);
  if (p.Ct) (void)hipFree(p.Ct);
  if (p.sq) (void)hipFree(p.sq);
  if (p.P) (void)hipFree(p.P);
  if (p.Pt) (void)hipFree(p.Pt);
  p = DctPlan();
}

// Autotuner of the folded stages, per grid size: over a table of the jobs of
// that size, every candidate shape that divides n/2 runs twice untimed, then
// three times between events on stream s; the fastest wins and is assigned
// to those jobs.  Launches are ungated and work on the jobs' scratch fields
// (values irrelevant to timing).  Results are cached per device, grid size and
// number of jobs of that size for the life of the process.
void dct_tune(JobSet& js, int device, hipStream_t s) {
  if (!js.fold) return;
  static std::mutex mu;
  static std::map<std::array<int, 3>, std::array<int, 4>> cache;
  std::map<int, std::vector<int>> by_size;
  for (int p = 0; p < js.count; ++p) by_size[js.host[p].f.n].push_back(p);
  for (const auto& kv : by_size) {
    const int n = kv.first;
    const std::array<int, 3> key = {device, n, int(kv.second.size())};
    std::array<int, 4> best{};
    bool hit = false;
    {
      std::lock_guard<std::mutex> lock(mu);
      auto it = cache.find(key);
      if (it != cache.end()) {
        best = it->second;
        hit = true;
      }
    }
    if (!hit) {
      std::vector<Job> jobs;
      for (int p : kv.second) jobs.push_back(js.host[p]);
      JobSet t;
      jobset_create(t, jobs, true);
      for (int st = 0; st < 4; ++st)
        best[st] = tune_pick(NFOLD, [&](int c) { return fold_cfg_valid(n, c); },
                             [&](int c) { FOLD_TABLE[st][c](t, false, false, s); }, s);
      OTM_HIP(hipStreamSynchronize(s));
      jobset_destroy(t);
      std::lock_guard<std::mutex> lock(mu);
      cache[key] = best;
    }
    for (int p : kv.second) js.jcfg[p] = best;
  }
  jobset_assign(js);
}

// one GEMM stage as the solve runs it, ungated, for kernel timing: folded F1
// (n^3 flops) when the table folds, else the unfolded T3 = C^T Y (2 n^3)
void launch_dct_gemm_stage(const JobSet& js, const DctPlan* p, hipStream_t s) {
  if (js.fold) {
    for (int c = 0; c < NFOLD; ++c) FOLD_TABLE[F1][c](js, false, false, s);
  } else {
    const Fields& f = js.host[0].f;
    gemm<G_PLAIN>(p->Ct, f.tmp1, f.tmp0, p->n, nullptr, p->sq, nullptr, s);
  }
}

// d_hat = C^T [(C R' C^T) / lambda] C of every job, R' = r - (sum r) e0.
// Every GEMM is skipped for jobs whose state is done; the last stage leaves
// the |d_hat|^2 partials in SITE_DNORM.  Without folding (n % 64 != 0) the
// table holds one job and p is its plan.
void launch_dct_solve(const JobSet& js, const DctPlan* p, hipStream_t s) {
  if (js.fold) {
    // one launch per tile shape in use (shapes with no job return at once)
    for (int st = 0; st < 4; ++st)
      for (int c = 0; c < NFOLD; ++c) FOLD_TABLE[st][c](js, true, st == F1, s);
    return;
  }
  const Fields& f = js.host[0].f;
  const int n = p->n;
  SolveState* st = f.state;
  gemm<G_RHS>(p->C, f.r, f.tmp0, n, st, p->sq, nullptr, s);
  gemm<G_SPECTRAL>(f.tmp0, p->Ct, f.tmp1, n, st, p->sq, nullptr, s);
  gemm<G_PLAIN>(p->Ct, f.tmp1, f.tmp0, n, st, p->sq, nullptr, s);
  gemm<G_FINAL>(f.tmp0, p->C, f.dhat, n, st, p->sq, f.site(SITE_DNORM), s);
}

}  // namespace otm
