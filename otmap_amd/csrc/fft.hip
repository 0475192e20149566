// FP64 Laplacian solve by fast DCTs (the transform route OTM_TRANSFORM_FFT).
//
// The same exact solve as the GEMM routes (dct.hip): with the orthonormal
// DCT-II C of size n, the pinned Neumann Laplacian of the reference
// (otlib/otsolver_2dgrid.cpp:284-321, solve and mean removal :174-183) is
//   d_hat = C^T [ (C R' C^T) / lambda ] C,   R' = R - (sum r) e_0 e_0^T,
// lambda_kl = 2 (sin^2(pi (k-l)/2n) + sin^2(pi (k+l)/2n)), d_hat(0,0)-mode = 0.
// Here every 1-D DCT runs as an n-point complex FFT, O(n^2 log n) work per
// solve instead of the 4 n^3 of the folded GEMMs, so the solve is bound by
// memory traffic (three passes of one read and one write of the n x n field).
//
// 1-D DCT-II of a real row x (Makhoul): v[m] = x[2m], v[n-1-m] = x[2m+1];
// V = FFT(v); X[k] = s_k Re(e^{-i pi k/2n} V[k]), s_0 = sqrt(1/n), s_k =
// sqrt(2/n).  Two real rows a, b share one complex FFT of z = v_a + i v_b:
// V_a[k] = (Z[k] + conj Z[n-k]) / 2, V_b[k] = (Z[k] - conj Z[n-k]) / 2i.
// DCT-III (the inverse): with U[k] = X[k] / s_k (U[n] = 0),
// V[k] = e^{i pi k/2n} (U[k] - i U[n-k]); v = IFFT(V); x[2m] = v[m],
// x[2m+1] = v[n-1-m]; again two rows per complex sequence.
//
// Passes (workgroups of 2 rows per complex sequence, packed, for n <= 1024;
// of 4 rows, two sequences, above; for n = 512 one wave per 2 rows in the
// register form, for n = 1024 one wave per row in the row form below):
//   1  rows of R' -> DCT-II along j, written transposed (B^T, 16- or 32-byte runs)
//   2  rows of B^T (columns of B) -> DCT-II along i, / lambda, DCT-III along
//      i, written transposed back (spatial i, frequency l along j)
//   3  rows -> DCT-III along j = d_hat, plus the |d_hat|^2 partials
// The FFT is a self-sorting (Stockham) mixed-radix FFT, radices 8, 4, 2, 3, in
// two LDS buffers; the stage twiddles (stage-major, n - 1 entries) are staged
// in LDS.
// Supported sizes: n = 2^a 3^b, 8 <= n <= 2048, n % 4 == 0.

#include <atomic>
#include <cfloat>
#include <cmath>
#include <vector>

#include "otm_internal.h"
#include "reduce.h"

// The transform's arithmetic follows no reference operation order (the
// reference solves by a sparse factorisation), so its multiply-adds contract
// to fused FMAs.
#pragma clang fp contract(fast)

namespace otm {

bool fft_supported(int n) {
  if (n < 8 || n > FFT_MAX_N || n % 4 != 0) return false;
  int m = n;
  while (m % 2 == 0) m /= 2;
  while (m % 3 == 0) m /= 3;
  return m == 1;
}

// Radix of the FFT stage that follows span L with `rest` points left to
// combine: 8 first, then 4, 2 and 3 (the device's fft_seq takes the same).
static int stage_radix(int rest) { return rest % 8 == 0 ? 8 : rest % 4 == 0 ? 4 : rest % 2 == 0 ? 2 : 3; }

void fft_plan_create(double*& table, int n) {
  table = nullptr;
  if (!fft_supported(n)) return;
  // Twiddles by stage: the stage of radix R after span L (L = product of the
  // earlier radices) uses exp(-2 pi i j r / (L R)), j < L, 1 <= r < R, at
  // entry (L - 1) + (r - 1) L + j -- consecutive j consecutive, so the lanes
  // of a stage read consecutive LDS words; the stages fill entries 0 .. n-2.
  // Then the DCT phases exp(-i pi k / 2n), k < n.
  std::vector<double> h(4 * size_t(n), 0.);
  const long double pi = 3.141592653589793238462643383279502884L;
  auto put = [&](size_t e, long double a) {
    h[2 * e] = double(cosl(a));
    h[2 * e + 1] = double(-sinl(a));
  };
  if (fft_row_form(n)) {
    // Row form.  Its m = n/2-point complex FFT (radices 4, 4, 4, 4 for m =
    // 256; 8, 8, 8 for 512; 16, 16, 4 for 1024): the twiddles of the stages
    // before the last by stage as above (entries 0 .. L - 2, L the span
    // before the last stage), then the quarter table
    // exp(-2 pi i q / m), q < m/4, from which the last stage takes its
    // twiddle exp(-2 pi i r j / m) as (-i)^(rj div m/4) times entry rj mod
    // m/4 (the rotation is exact).  Then the quarter table exp(-2 pi i q / n),
    // q < n/4, of the real-FFT split.  Last, the DCT
    // phases of k <= n/2 (phase n - k is -i conj(phase k): the real and
    // imaginary parts swapped and negated).
    const int m = n / 2, R = m == 256 ? 4 : m == 512 ? 8 : 16;
    int L = 1;
    for (; L * R < m; L *= R)
      for (int r = 1; r < R; ++r)
        for (int j = 0; j < L; ++j) put(size_t(L - 1 + (r - 1) * L + j), 2.0L * pi * (j * r) / (long double)(L * R));
    size_t e = size_t(L - 1);  // L: the span before the last stage (m / its radix)
    for (int q = 0; q < m / 4; ++q) put(e++, 2.0L * pi * q / (long double)m);
    for (int q = 0; q < n / 4; ++q) put(e++, 2.0L * pi * q / (long double)n);
    for (int t = 0; t <= n / 2; ++t) put(e++, pi * t / (2.0L * n));
  } else {
    for (int rest = n, L = 1; rest > 1;) {
      const int R = stage_radix(rest);
      for (int r = 1; r < R; ++r)
        for (int j = 0; j < L; ++j) put(size_t(L - 1 + (r - 1) * L + j), 2.0L * pi * (j * r) / (long double)(L * R));
      L *= R;
      rest /= R;
    }
    for (int t = 0; t < n; ++t) put(size_t(n) + size_t(t), pi * t / (2.0L * n));
  }
  OTM_HIP(hipMalloc(&table, h.size() * sizeof(double)));
  OTM_HIP(hipMemcpy(table, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
}

namespace {

typedef double cd __attribute__((ext_vector_type(2)));  // complex (re, im)

// A workgroup of a pass holds FS complex sequences (2 FS grid rows) and has
// 256 FS threads; thread t owns the frequencies / columns k = t + u 256 FS,
// u < KPT, of every sequence: FS = 1 serves n <= 1024, FS = 2 n <= 2048.
constexpr int KPT = 4;
static_assert(FFT_SPLIT_N == 256 * KPT && FFT_MAX_N == 512 * KPT, "four columns per thread");
constexpr int fft_lds(int fs) { return 16 * (2 * fs + 1); }  // LDS bytes per unit of n

__device__ inline cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline cd conjc(cd a) { return {a.x, -a.y}; }
__device__ inline cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline cd csub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
// multiply by -i (sign < 0) or +i (sign > 0)
__device__ inline cd mul_i(cd a, int sign) { return sign < 0 ? cd{a.y, -a.x} : cd{-a.y, a.x}; }

// DFT of 4 points with w4 = e^{SIGN 2 pi i / 4} = SIGN i
template <int SIGN>
__device__ inline void dft4(cd a0, cd a1, cd a2, cd a3, cd& x0, cd& x1, cd& x2, cd& x3) {
  const cd s02 = cadd(a0, a2), d02 = csub(a0, a2);
  const cd s13 = cadd(a1, a3), d13 = mul_i(csub(a1, a3), SIGN);
  x0 = cadd(s02, s13);
  x1 = cadd(d02, d13);
  x2 = csub(s02, s13);
  x3 = csub(d02, d13);
}

// DFT of 8 points: X[2q] = DFT4(a_r + a_{r+4})[q], X[2q+1] = DFT4((a_r -
// a_{r+4}) w8^r)[q], w8 = e^{SIGN 2 pi i / 8}: w8^1 = (1 + SIGN i)/sqrt 2,
// w8^2 = SIGN i, w8^3 = (-1 + SIGN i)/sqrt 2
template <int SIGN, int RA>
__device__ inline void dft8(const cd (&a)[RA], cd (&c)[RA]) {
  static_assert(RA >= 8, "eight points");
  const double h = 0.70710678118654752440;  // 1 / sqrt(2)
  cd s[4], d[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    s[r] = cadd(a[r], a[r + 4]);
    d[r] = csub(a[r], a[r + 4]);
  }
  d[1] = cd{h * (d[1].x - SIGN * d[1].y), h * (d[1].y + SIGN * d[1].x)};
  d[2] = mul_i(d[2], SIGN);
  d[3] = cd{h * (-d[3].x - SIGN * d[3].y), h * (-d[3].y + SIGN * d[3].x)};
  dft4<SIGN>(s[0], s[1], s[2], s[3], c[0], c[2], c[4], c[6]);
  dft4<SIGN>(d[0], d[1], d[2], d[3], c[1], c[3], c[5], c[7]);
}

// One radix-R stage of the self-sorting FFT of the FS sequences held in src
// (seq q at src[q N .. q N + N)), written to dst: butterfly b of sequence q,
// j = b % L, m = b / L, takes src[b + r N/R] times w^(j r) (w = e^{SIGN 2 pi
// i / (L R)}, tw[(L - 1) + (r - 1) L + j] in the stage-major table), and its
// DFT_R output q' goes to dst[m L R + j + q' L].  The first stage (L = 1)
// has no twiddles; later stages multiply every input by its twiddle, also
// w^0 = (1, 0), which is exact, so the butterfly has no branches and all its
// LDS reads are issued together.
template <int R, int SIGN>
__device__ inline void fft_butterfly(const cd* __restrict__ s, cd* __restrict__ dst, const cd* __restrict__ tw, int NR,
                                     int L, int b, int j, int m) {
  cd a[R], w[R];
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = s[b + r * NR];
  if (L > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) w[r] = tw[(L - 1) + (r - 1) * L + j];
#pragma unroll
    for (int r = 1; r < R; ++r) a[r] = cmul(a[r], SIGN < 0 ? w[r] : conjc(w[r]));
  }
  cd c[R];
  if constexpr (R == 2) {
    c[0] = cadd(a[0], a[1]);
    c[1] = csub(a[0], a[1]);
  } else if constexpr (R == 4) {
    dft4<SIGN>(a[0], a[1], a[2], a[3], c[0], c[1], c[2], c[3]);
  } else if constexpr (R == 8) {
    dft8<SIGN>(a, c);
  } else {  // R == 3: w3 = e^{SIGN 2 pi i / 3} = (-1/2, SIGN sqrt(3)/2)
    const double h3 = 0.86602540378443864676;  // sqrt(3) / 2
    const cd s12 = cadd(a[1], a[2]), d12 = csub(a[1], a[2]);
    const cd t = {a[0].x - 0.5 * s12.x, a[0].y - 0.5 * s12.y};
    const cd u = mul_i(cd{h3 * d12.x, h3 * d12.y}, SIGN);  // SIGN i sqrt(3)/2 (a1 - a2)
    c[0] = cadd(a[0], s12);
    c[1] = cadd(t, u);
    c[2] = csub(t, u);
  }
  cd* d = dst + m * L * R + j;
#pragma unroll
  for (int r = 0; r < R; ++r) d[r * L] = c[r];
}

template <int FS, int R, int SIGN>
__device__ inline void fft_stage(const cd* __restrict__ src, cd* __restrict__ dst, const cd* __restrict__ tw, int N,
                                 int L) {
  constexpr int FT = 256 * FS;
  const int NR = N / R;
  const bool pow2 = (L & (L - 1)) == 0;
  const int lshift = __ffs(L) - 1;
  for (int b2 = threadIdx.x; b2 < FS * NR; b2 += FT) {
    const int q = FS == 1 ? 0 : b2 >= NR ? 1 : 0;
    const int b = b2 - q * NR;
    const int j = pow2 ? (b & (L - 1)) : b % L, m = pow2 ? (b >> lshift) : b / L;
    fft_butterfly<R, SIGN>(src + q * N, dst + q * N, tw, NR, L, b, j, m);
  }
}

// The same stage in the packed form: the thread's own sequence (src, dst
// already offset to it), butterflies tt, tt + tps, ...
template <int R, int SIGN>
__device__ inline void fft_stage_p(const cd* __restrict__ src, cd* __restrict__ dst, const cd* __restrict__ tw, int N,
                                   int L, int tt, int tps) {
  const int NR = N / R;
  const bool pow2 = (L & (L - 1)) == 0;
  const int lshift = __ffs(L) - 1;
  for (int b = tt; b < NR; b += tps) {
    const int j = pow2 ? (b & (L - 1)) : b % L, m = pow2 ? (b >> lshift) : b / L;
    fft_butterfly<R, SIGN>(src, dst, tw, NR, L, b, j, m);
  }
}

// FFT of every sequence of a packed workgroup: sequence q at a[q N ..] (b
// scratch); threads of no sequence (act false) only keep the barriers.
// Returns the buffer that holds the result (radices as fft_seq).
template <int SIGN>
__device__ inline cd* fft_seq_p(cd* a, cd* b, const cd* tw, int N, int qoff, int tt, int tps, bool act) {
  int rest = N, L = 1;
  while (rest > 1) {
    const int R = rest % 8 == 0 ? 8 : rest % 4 == 0 ? 4 : rest % 2 == 0 ? 2 : 3;
    if (act) {
      if (R == 8) fft_stage_p<8, SIGN>(a + qoff, b + qoff, tw, N, L, tt, tps);
      else if (R == 4) fft_stage_p<4, SIGN>(a + qoff, b + qoff, tw, N, L, tt, tps);
      else if (R == 2) fft_stage_p<2, SIGN>(a + qoff, b + qoff, tw, N, L, tt, tps);
      else fft_stage_p<3, SIGN>(a + qoff, b + qoff, tw, N, L, tt, tps);
    }
    L *= R;
    rest /= R;
    __syncthreads();
    cd* t = a;
    a = b;
    b = t;
  }
  return a;
}

// FFT of the FS sequences of `a` (length N each), `b` scratch; returns the
// buffer that holds the result.  Radices 8 first, then 4, 2 and 3 (fewer
// stages: every stage ends in a workgroup barrier), as stage_radix on the
// host, which lays out the stage twiddles in this order.
template <int FS, int SIGN>
__device__ inline cd* fft_seq(cd* a, cd* b, const cd* tw, int N) {
  int rest = N, L = 1;
  while (rest > 1) {
    if (rest % 8 == 0) {
      fft_stage<FS, 8, SIGN>(a, b, tw, N, L);
      L *= 8;
      rest /= 8;
    } else if (rest % 4 == 0) {
      fft_stage<FS, 4, SIGN>(a, b, tw, N, L);
      L *= 4;
      rest /= 4;
    } else if (rest % 2 == 0) {
      fft_stage<FS, 2, SIGN>(a, b, tw, N, L);
      L *= 2;
      rest /= 2;
    } else {
      fft_stage<FS, 3, SIGN>(a, b, tw, N, L);
      L *= 3;
      rest /= 3;
    }
    __syncthreads();
    cd* t = a;
    a = b;
    b = t;
  }
  return a;
}

// position of x[c] in the permuted sequence v (v[m] = x[2m], v[N-1-m] = x[2m+1])
__device__ inline int perm_pos(int c, int N) { return (c & 1) ? N - 1 - (c >> 1) : (c >> 1); }

// orthonormal DCT scale s_k and its reciprocal (k == 0 or not)
struct Scales {
  double s0, s1, r0, r1;
  __device__ explicit Scales(int N)
      : s0(sqrt(1.0 / N)), s1(sqrt(2.0 / N)), r0(sqrt(double(N))), r1(sqrt(0.5 * N)) {}
  __device__ double s(int k) const { return k == 0 ? s0 : s1; }
  __device__ double r(int k) const { return k == 0 ? r0 : r1; }
};

// forward post-processing of one sequence at frequency k from Z[k] and
// Z[N-k] (Z[0] for k = 0), phase p = e^{-i pi k/2N}, scale s = s_k: the
// DCT-II coefficients of its two rows (real part row 2q, imaginary part row 2q+1)
__device__ inline void dct2_post_pair(cd Z, cd Zn, cd p, double s, double& xa, double& xb) {
  const cd Zc = conjc(Zn);
  const cd va = {0.5 * (Z.x + Zc.x), 0.5 * (Z.y + Zc.y)};
  const cd dd = csub(Z, Zc);
  const cd vb = {0.5 * dd.y, -0.5 * dd.x};  // (Z - Zc) / 2i
  xa = s * (p.x * va.x - p.y * va.y);
  xb = s * (p.x * vb.x - p.y * vb.y);
}
__device__ inline void dct2_post(const cd* z, cd p, const Scales& sc, int N, int k, double& xa, double& xb) {
  dct2_post_pair(z[k], z[k == 0 ? 0 : N - k], p, sc.s(k), xa, xb);
}

// inverse pre-processing at frequency k from the scaled coefficients U = X /
// s_k of two rows (ua = U_a[k], uan = U_a[N-k], 0 for k = 0; likewise b):
// Z[k] = V_a[k] + i V_b[k], V[k] = e^{i pi k/2N} (U[k] - i U[N-k])
__device__ inline cd dct3_pre_u(double ua, double uan, double ub, double ubn, cd p) {
  const cd e = conjc(p);  // e^{i pi k / 2N}
  const cd va = cmul(e, cd{ua, -uan}), vb = cmul(e, cd{ub, -ubn});
  return {va.x - vb.y, va.y + vb.x};  // V_a + i V_b
}
// the same from the DCT-II coefficients (a = row 2q, b = row 2q+1; xan = X_a[N-k])
__device__ inline cd dct3_pre(double xa, double xan, double xb, double xbn, cd p, const Scales& sc, int k) {
  const double r = sc.r(k);
  const double uan = k == 0 ? 0. : xan * sc.r1, ubn = k == 0 ? 0. : xbn * sc.r1;
  return dct3_pre_u(xa * r, uan, xb * r, ubn, p);
}

struct FftLoc {
  int p, lt, nb;
};
// The dispatcher deals consecutive workgroups to the 8 XCDs in turn.  The
// first 8 floor(G/8) workgroups are renumbered so that XCD x takes the
// consecutive row groups x G/8 .. (x+1) G/8 - 1: the 16- or 32-byte column
// segments of the transposed writes of neighbouring row groups then meet in
// the same L2 and leave it as whole lines.
__device__ inline FftLoc fft_locate(const int* __restrict__ starts, int count) {
  const int g = int(gridDim.x), full = g & ~7;
  int b = int(blockIdx.x);
  if (b < full) b = (b & 7) * (full >> 3) + (b >> 3);
  FftLoc L;
  L.p = count == 1 ? 0 : job_of(starts, count, b);
  L.lt = count == 1 ? b : b - starts[L.p];
  L.nb = count == 1 ? g : starts[L.p + 1] - starts[L.p];
  return L;
}

// LDS: buffers A and B of FS sequences x N complex each, then the twiddles
// (N complex): fft_lds(FS) N bytes.  Every global read of a pass is issued
// at its start.

struct Prefetch {
  cd tw[KPT];  // twiddles t + u FT
  cd ph[KPT];  // DCT phases of k = t + u FT
};
template <int FS>
__device__ inline void prefetch_tables(Prefetch& P, const double* table, int N) {
  constexpr int FT = 256 * FS;
  const auto tw = gp(reinterpret_cast<const cd*>(table));
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int k = threadIdx.x + u * FT;
    P.tw[u] = k < N ? tw[k] : cd{0., 0.};
    P.ph[u] = k < N ? tw[N + k] : cd{0., 0.};
  }
}
template <int FS>
__device__ inline void store_twiddles(const Prefetch& P, cd* twl, int N) {
  constexpr int FT = 256 * FS;
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int k = threadIdx.x + u * FT;
    if (k < N) twl[k] = P.tw[u];
  }
}

// rows r0 .. r0 + 2 FS - 1 of `in` (row length N) into registers, column k = t + u 256 FS
template <int FS>
__device__ inline void load_rows(double (&v)[2 * FS][KPT], const double* in_, int N, int r0) {
  constexpr int FT = 256 * FS;
  const auto in = gp(in_);
#pragma unroll
  for (int rr = 0; rr < 2 * FS; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int c = threadIdx.x + u * FT;
      v[rr][u] = c < N ? in[long(r0 + rr) * N + c] : 0.;
    }
}
// the rows into the complex sequences of A, permuted (v[m] = x[2m], v[N-1-m] = x[2m+1])
template <int FS>
__device__ inline void store_permuted(const double (&v)[2 * FS][KPT], cd* A, int N) {
  constexpr int FT = 256 * FS;
#pragma unroll
  for (int rr = 0; rr < 2 * FS; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int c = threadIdx.x + u * FT;
      if (c < N) reinterpret_cast<double*>(A + (rr >> 1) * N + perm_pos(c, N))[rr & 1] = v[rr][u];
    }
}

// Pass 1: rows of R' (R minus (sum r) at (0,0) when rhs), DCT-II along the
// row, written transposed: out[k * n + row].
template <int FS>
__global__ __launch_bounds__(256 * FS) void k_fft_rows_fwd(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                    int count, bool gate, bool rhs) {
  constexpr int FT = 256 * FS, ROWS = 2 * FS;
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  // one job: its reads are issued before the gate is read (the two latencies
  // overlap); in a batch the gate comes first, so converged jobs read nothing
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const int N = J.f.n, r0 = ROWS * L.lt;
  cd* A = fsm;
  cd* B = fsm + FS * N;
  cd* twl = fsm + 2 * FS * N;
  Prefetch P;
  prefetch_tables<FS>(P, J.fft, N);
  double v[ROWS][KPT];
  load_rows<FS>(v, J.f.r, N, r0);
  if (rhs && r0 == 0 && threadIdx.x == 0) v[0][0] -= st->sum_r;  // R' = R - (sum r) e0 e0^T
  if (gate && early && st->done) return;
  store_twiddles<FS>(P, twl, N);
  store_permuted<FS>(v, A, N);
  __syncthreads();
  const cd* z = fft_seq<FS, -1>(A, B, twl, N);
  const Scales sc(N);
  const auto out = gp(J.f.tmp0);
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int k = threadIdx.x + u * FT;
    if (k >= N) break;
    auto o = reinterpret_cast<OTM_GLOBAL cd*>(out + long(k) * N + r0);
#pragma unroll
    for (int q = 0; q < FS; ++q) {
      double x0, x1;
      dct2_post(z + q * N, P.ph[u], sc, N, k, x0, x1);
      o[q] = cd{x0, x1};
    }
  }
}

// Pass 2: rows of B^T (row l: DCT frequency l along j), DCT-II along i,
// divided by lambda(k, l) (the (0,0) mode set to 0), DCT-III along i,
// written transposed: out[i * n + l].
template <int FS>
__global__ __launch_bounds__(256 * FS) void k_fft_cols(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                int count, bool gate) {
  constexpr int FT = 256 * FS, ROWS = 2 * FS;
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  // one job: its reads are issued before the gate is read (the two latencies
  // overlap); in a batch the gate comes first, so converged jobs read nothing
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const int N = J.f.n, l0 = ROWS * L.lt;
  cd* A = fsm;
  cd* B = fsm + FS * N;
  cd* twl = fsm + 2 * FS * N;
  Prefetch P;
  prefetch_tables<FS>(P, J.fft, N);
  double v[ROWS][KPT];
  load_rows<FS>(v, J.f.tmp0, N, l0);
  // eigenvalue halves sin^2(pi |k-l| / 2N) + sin^2(pi (k+l) / 2N) of this thread's (k, l)
  const auto sq = gp(J.sq);
  double lam[ROWS][KPT];
#pragma unroll
  for (int rr = 0; rr < ROWS; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = threadIdx.x + u * FT, l = l0 + rr;
      const int dm = k > l ? k - l : l - k;
      lam[rr][u] = k < N ? sq[dm] + sq[k + l] : 1.;
    }
  if (gate && early && st->done) return;
  store_twiddles<FS>(P, twl, N);
  store_permuted<FS>(v, A, N);
  __syncthreads();
  cd* z = fft_seq<FS, -1>(A, B, twl, N);
  cd* other = z == A ? B : A;
  const Scales sc(N);
  // spectrum / lambda as ROWS real rows in the free buffer
  double* X = reinterpret_cast<double*>(other);
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int k = threadIdx.x + u * FT;
    if (k >= N) break;
    double x[ROWS];
#pragma unroll
    for (int q = 0; q < FS; ++q) dct2_post(z + q * N, P.ph[u], sc, N, k, x[2 * q], x[2 * q + 1]);
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr)
      X[rr * N + k] = (k == 0 && l0 + rr == 0) ? 0. : x[rr] / (2. * lam[rr][u]);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int k = threadIdx.x + u * FT;
    if (k >= N) break;
    const int kn = k == 0 ? 0 : N - k;
#pragma unroll
    for (int q = 0; q < FS; ++q) {
      const double* Xa = X + 2 * q * N;
      const double* Xb = Xa + N;
      z[q * N + k] = dct3_pre(Xa[k], Xa[kn], Xb[k], Xb[kn], P.ph[u], sc, k);
    }
  }
  __syncthreads();
  const cd* y = fft_seq<FS, 1>(z, other, twl, N);
  const auto out = gp(J.f.tmp1);
  const double inv = 1.0 / N;
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int i = threadIdx.x + u * FT;
    if (i >= N) break;
    const int m = perm_pos(i, N);
    auto o = reinterpret_cast<OTM_GLOBAL cd*>(out + long(i) * N + l0);
#pragma unroll
    for (int q = 0; q < FS; ++q) {
      const cd a = y[q * N + m];
      o[q] = cd{a.x * inv, a.y * inv};
    }
  }
}

// Pass 3: rows (spatial i, frequency l along j), DCT-III along j = d_hat,
// and one |d_hat|^2 partial per workgroup.
template <int FS>
__global__ __launch_bounds__(256 * FS) void k_fft_rows_inv(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                    int count, bool gate) {
  constexpr int FT = 256 * FS, ROWS = 2 * FS;
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  // one job: its reads are issued before the gate is read (the two latencies
  // overlap); in a batch the gate comes first, so converged jobs read nothing
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const int N = J.f.n, r0 = ROWS * L.lt;
  cd* A = fsm;
  cd* B = fsm + FS * N;
  cd* twl = fsm + 2 * FS * N;
  Prefetch P;
  prefetch_tables<FS>(P, J.fft, N);
  // X[row][k] and X[row][N-k] of the ROWS rows
  const auto in = gp((const double*)J.f.tmp1) + long(r0) * N;
  double xk[ROWS][KPT], xn[ROWS][KPT];
#pragma unroll
  for (int rr = 0; rr < ROWS; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = threadIdx.x + u * FT;
      const int kn = k == 0 ? 0 : N - k;
      xk[rr][u] = k < N ? in[rr * N + k] : 0.;
      xn[rr][u] = k < N ? in[rr * N + kn] : 0.;
    }
  if (gate && early && st->done) return;
  store_twiddles<FS>(P, twl, N);
  const Scales sc(N);
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int k = threadIdx.x + u * FT;
    if (k >= N) break;
#pragma unroll
    for (int q = 0; q < FS; ++q)
      A[q * N + k] = dct3_pre(xk[2 * q][u], xn[2 * q][u], xk[2 * q + 1][u], xn[2 * q + 1][u], P.ph[u], sc, k);
  }
  __syncthreads();
  const cd* y = fft_seq<FS, 1>(A, B, twl, N);
  const auto out = gp(J.f.dhat);
  const double inv = 1.0 / N;
  double nrm[1] = {0.};
#pragma unroll
  for (int rr = 0; rr < ROWS; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int c = threadIdx.x + u * FT;
      if (c >= N) break;
      const cd a = y[(rr >> 1) * N + perm_pos(c, N)];
      const double x = ((rr & 1) ? a.y : a.x) * inv;
      out[long(r0 + rr) * N + c] = x;
      nrm[0] += x * x;
    }
  // workgroup partial: DPP wave sums, then the wave totals in a fixed order
  // through LDS, in the buffer that does not hold y (the last FFT stage read
  // it before its closing barrier), so the kernel has no static LDS
  wave_sum<1>(nrm);
  double* red = reinterpret_cast<double*>(y == A ? B : A);
  if ((threadIdx.x & 63) == 63) red[threadIdx.x >> 6] = nrm[0];
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.;
    for (int w = 0; w < FT / 64; w += 4) x += (red[w] + red[w + 1]) + (red[w + 2] + red[w + 3]);
    gp(J.f.site(SITE_DNORM))[L.lt] = x;
  }
}

// ---------------------------------------------------------------------------
// Register FFT (the row form below runs it per grid row): one wave per
// complex sequence of N = 64 E points, lane t holding the E points t + 64 v
// (v < E) in registers.  A stage of radix R after span L gives lane t the
// butterflies b = t + 64 u (u < E / R), whose inputs b + r N / R are its own
// slots u + r E / R; its outputs go through the wave's LDS exchange buffer
// back into the same layout, except in the last stage (L R = N), where they
// land in the lane's own slots.  Radices 8, 8, 8 (N = 512) or 4, 4, 4, 4
// (N = 256): two or three exchanges per FFT instead of a workgroup barrier
// after each stage.  The exchange buffer's positions are padded, p + p / 16,
// so the lanes of a 16-lane group write distinct banks.
// ---------------------------------------------------------------------------
template <int E>
struct RegRadix;
template <>
struct RegRadix<8> {  // 512 points: 8 x 8 x 8
  static constexpr int R = 8, LAST = 8;
};
template <>
struct RegRadix<4> {  // 256 points: 4 x 4 x 4 x 4
  static constexpr int R = 4, LAST = 4;
};
template <>
struct RegRadix<16> {  // 1024 points: 16 x 16 x 4
  static constexpr int R = 16, LAST = 4;
};

// The register FFT's twiddles (fft_plan_create): the stage-major twiddles
// of the stages before the last (entries 0 .. Q0 - 1) and the quarter table
// e^{-2 pi i m / N}, m < N/4, for the last stage (PH0: the next entry).
template <int E>
struct RegTab {
  static constexpr int N = 64 * E;
  static constexpr int Q0 = N / RegRadix<E>::LAST - 1;
  static constexpr int PH0 = Q0 + N / 4;
};
// the last stage's twiddle e^{-2 pi i m / N}, m < N: (-i)^(m div N/4) times
// the quarter-table entry m mod N/4 (the rotation is exact)
template <int E>
__device__ inline cd reg_twiddle(const cd* tab, int m) {
  constexpr int Q = 16 * E;  // N / 4
  const cd q = tab[RegTab<E>::Q0 + (m & (Q - 1))];
  const int s = m / Q;
  cd w = (s & 1) ? cd{q.y, -q.x} : q;
  if (s & 2) w = cd{-w.x, -w.y};
  return w;
}
__device__ inline int xpad(int p) { return p + (p >> 4); }

// DFT of 16 points as 4 x 4: DFT4 over n1 of a[n2 + 4 n1], times w16^(n2 k1),
// DFT4 over n2; X[k1 + 4 k2] lands in c[k1 + 4 k2]
template <int SIGN>
__device__ inline cd w16mul(cd y, int m) {
  const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;  // cos, sin(pi/8)
  const double h = 0.70710678118654752440;
  switch (m) {
    case 1: return cmul(y, cd{c1, SIGN * s1});
    case 2: return cmul(y, cd{h, SIGN * h});
    case 3: return cmul(y, cd{s1, SIGN * c1});
    case 4: return mul_i(y, SIGN);
    case 6: return cmul(y, cd{-h, SIGN * h});
    case 9: return cmul(y, cd{-c1, -SIGN * s1});
    default: return y;
  }
}
template <int SIGN>
__device__ inline void dft16(const cd (&a)[16], cd (&c)[16]) {
  cd y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2)
    dft4<SIGN>(a[n2], a[n2 + 4], a[n2 + 8], a[n2 + 12], y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) y[n2][k1] = w16mul<SIGN>(y[n2][k1], n2 * k1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<SIGN>(y[0][k1], y[1][k1], y[2][k1], y[3][k1], c[k1], c[k1 + 4], c[k1 + 8], c[k1 + 12]);
}


template <int E, int R, int L, int SIGN>
__device__ inline void rstage(cd (&x)[E], cd* xb, const cd* tw, int t) {
  constexpr int N = 64 * E, U = E / R;
  constexpr bool LAST = L * R == N;
  static_assert(E % R == 0 && (!LAST || L == 64 * U), "lane slots");
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int b = t + 64 * u;
    const int j = b & (L - 1), m = b / L;
    cd a[R], c[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = x[u + r * U];
    if constexpr (L > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cd w = LAST ? reg_twiddle<E>(tw, r * j) : tw[(L - 1) + (r - 1) * L + j];
        a[r] = cmul(a[r], SIGN < 0 ? w : conjc(w));
      }
    }
    static_assert(R == 4 || R == 8 || R == 16, "radix 4, 8 or 16");
    if constexpr (R == 4) dft4<SIGN>(a[0], a[1], a[2], a[3], c[0], c[1], c[2], c[3]);
    else if constexpr (R == 8) dft8<SIGN>(a, c);
    else dft16<SIGN>(a, c);
    if constexpr (LAST) {
#pragma unroll
      for (int q = 0; q < R; ++q) x[u + q * U] = c[q];  // position b + q L = t + 64 (u + q U)
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) xb[xpad(m * L * R + j + q * L)] = c[q];
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
#pragma unroll
    for (int v = 0; v < E; ++v) x[v] = xb[xpad(t + 64 * v)];
    __syncthreads();
  }
}

// the stages of radix R after spans L = 1, R, R^2, ..., the last one of the
// radix that completes 64 E points
template <int E, int SIGN, int L = 1>
__device__ inline void reg_fft(cd (&x)[E], cd* xb, const cd* tw, int t) {
  constexpr int R = L * RegRadix<E>::R <= 64 * E ? RegRadix<E>::R : 64 * E / L;
  rstage<E, R, L, SIGN>(x, xb, tw, t);
  if constexpr (L * R < 64 * E) reg_fft<E, SIGN, L * R>(x, xb, tw, t);
}

// table entries per thread of a W-wave workgroup
constexpr int reg_tpt(int size, int w) { return (size + 64 * w - 1) / (64 * w); }
// the SIZE-entry table into registers (all 64 W threads), then into LDS
template <int SIZE, int W>
__device__ inline void tab_load(cd (&tv)[reg_tpt(SIZE, W)], const double* table) {
  const auto tg = gp(reinterpret_cast<const cd*>(table));
#pragma unroll
  for (int q = 0; q < reg_tpt(SIZE, W); ++q) {
    const int e = int(threadIdx.x) + 64 * W * q;
    tv[q] = e < SIZE ? tg[e] : cd{0., 0.};
  }
}
template <int SIZE, int W>
__device__ inline void tab_store(const cd (&tv)[reg_tpt(SIZE, W)], cd* tab) {
#pragma unroll
  for (int q = 0; q < reg_tpt(SIZE, W); ++q) {
    const int e = int(threadIdx.x) + 64 * W * q;
    if (e < SIZE) tab[e] = tv[q];
  }
}

// ---------------------------------------------------------------------------
// Row form (n = 1024 and 512, launch kinds KIND_FFTW8 and KIND_FFTW4): one
// wave per grid row, four rows per workgroup, so a single 1024^2 grid runs
// 1024 waves, one per SIMD, each with half the FFT work of a two-row wave.  The passes with a
// transposed output stage their four rows in LDS and write 32-byte runs
// (two lanes each), half the store line touches of the two-row form.  The DCT
// of a real row x needs V = FFT_n(v) of the real sequence v (Makhoul's order,
// v[j] = x[2j], v[n-1-j] = x[2j+1]); it comes from the n/2-point complex FFT
// Z of z[m] = v[2m] + i v[2m+1] (the register FFT, E = n/128 points per
// lane) and the split
//   E_k = (Z[k] + conj Z[n/2-k]) / 2,  O_k = (Z[k] - conj Z[n/2-k]) / 2i,
//   V[k] = E_k + w^k O_k,  V[k + n/2] = E_k - w^k O_k,  w = e^{-2 pi i/n}.
// The inverse (V Hermitian, v real) joins E_k = (V[k] + V[k+n/2]) / 2 and
// O_k = (V[k] - V[k+n/2]) conj(w^k) / 2 into Z[k] = E_k + i O_k, runs the
// inverse n/2-point FFT and reads z[m] / (n/2) = v[2m] + i v[2m+1].  Lane t
// holds m = t + 64 u (u < E) and the frequencies k = t + 64 u and k + n/2;
// frequency k pairs with n/2 - k (one exchange), which also holds n - k.
// z[m] holds x[4m], x[4m+2] for m < n/4 and x[2n-1-4m], x[2n-3-4m] above.
// ---------------------------------------------------------------------------
template <int E>
struct RowTab {
  static constexpr int N = 128 * E, M = N / 2;  // row length, FFT points (E per lane)
  static constexpr int WQ0 = RegTab<E>::PH0;  // after the n/2-point FFT's table: w^q, q < n/4
  static constexpr int PH0 = WQ0 + N / 4;     // DCT phases e^{-i pi k / 2n}, k <= n/2
  static constexpr int SIZE = PH0 + N / 2 + 1;
  static constexpr int XB = M + M / 16;  // exchange buffer per wave
  static constexpr int SN = N + 4;       // staging row stride (doubles), over the exchange buffers
  static constexpr size_t lds_bytes(int w) { return size_t(w * XB + SIZE) * sizeof(cd); }
  static_assert(SN <= 2 * XB, "a staging row fits its wave's exchange buffer");
  static_assert(RegTab<E>::N == M, "the row form's FFT is the register FFT of E points per lane");
};
// w^k = e^{-2 pi i k / n}, k < n/2
template <int E>
__device__ inline cd row_w(const cd* tab, int k) {
  constexpr int Q = RowTab<E>::N / 4;
  const cd q = tab[RowTab<E>::WQ0 + (k & (Q - 1))];
  return k >= Q ? cd{q.y, -q.x} : q;
}
// phases of k and k + n/2 (the latter as phase n/2 - k swapped and negated)
template <int E>
__device__ inline cd row_phase_lo(const cd* tab, int k) { return tab[RowTab<E>::PH0 + k]; }
template <int E>
__device__ inline cd row_phase_hi(const cd* tab, int k) {
  const cd p = tab[RowTab<E>::PH0 + RowTab<E>::M - k];
  return cd{-p.y, -p.x};
}
// the row's z in lane t, slot u (u < E/2: m < n/4): position of Re z[m] and of Im z[m] in the row
template <int E>
__device__ inline int row_c0(int u, int m) { return u < E / 2 ? 4 * m : 2 * RowTab<E>::N - 1 - 4 * m; }
template <int E>
__device__ inline int row_c1(int u, int m) { return u < E / 2 ? 4 * m + 2 : 2 * RowTab<E>::N - 3 - 4 * m; }
template <int E>
__device__ inline void row_load(cd (&x)[E], const double* in_, int row, int t) {
  const auto in = gp(in_) + long(row) * RowTab<E>::N;
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int m = t + 64 * u;
    x[u] = cd{in[row_c0<E>(u, m)], in[row_c1<E>(u, m)]};
  }
}
// DCT-II coefficients X[k], X[k + n/2] from Z[k] and Zp = Z[(n/2 - k) mod n/2]
__device__ inline void row_post(cd Z, cd Zp, cd w, cd plo, cd phi, double slo, double shi, double& xlo,
                                double& xhi) {
  const cd Zc = conjc(Zp);
  const cd e = {0.5 * (Z.x + Zc.x), 0.5 * (Z.y + Zc.y)};
  const cd d = csub(Z, Zc);
  const cd o = {0.5 * d.y, -0.5 * d.x};  // (Z - Zc) / 2i
  const cd wo = cmul(w, o);
  const cd vlo = cadd(e, wo), vhi = csub(e, wo);
  xlo = slo * (plo.x * vlo.x - plo.y * vlo.y);
  xhi = shi * (phi.x * vhi.x - phi.y * vhi.y);
}
// Z[k] = E_k + i O_k of the inverse from the scaled coefficients U = X / s_k:
// ulo = U[k], uhi = U[k + n/2], umk = U[n/2 - k], unk = U[n - k] (U[n] = 0);
// V[k] = e^{i pi k/2n} (U[k] - i U[n-k])
__device__ inline cd row_pre(double ulo, double uhi, double umk, double unk, cd w, cd plo, cd phi) {
  const cd vlo = cmul(conjc(plo), cd{ulo, -unk});
  const cd vhi = cmul(conjc(phi), cd{uhi, -umk});
  const cd e = {0.5 * (vlo.x + vhi.x), 0.5 * (vlo.y + vhi.y)};
  const cd o = cmul(cd{0.5 * (vlo.x - vhi.x), 0.5 * (vlo.y - vhi.y)}, conjc(w));
  return {e.x - o.y, e.y + o.x};
}
// the wave's lane, row, exchange buffer and the workgroup's table
template <int E, int W>
struct RowWave {
  int t, row;
  cd* xb;
  cd* tab;
  __device__ RowWave(cd* fsm, int lt) {
    const int wv = int(threadIdx.x) >> 6;
    t = int(threadIdx.x) & 63;
    row = W * lt + wv;
    xb = fsm + wv * RowTab<E>::XB;
    tab = fsm + W * RowTab<E>::XB;
  }
  // the wave's staging row over the exchange buffers (row stride RowTab::SN)
  __device__ double* stage(cd* fsm) const {
    return reinterpret_cast<double*>(fsm) + (int(threadIdx.x) >> 6) * RowTab<E>::SN;
  }
};
// The transposed write of a pass: the W rows of the workgroup (first row
// W lt) sit in the staging rows; output row k takes them as one W-value run
// at column W lt.  Lane pairs write the two 16-byte halves of a run, so a
// store instruction touches 32 lines instead of 64.
template <int E, int W>
__device__ inline void row_store_transposed(const cd* fsm, double* out_, int lt) {
  static_assert(W == 4, "two lanes per four-row run");
  constexpr int N = RowTab<E>::N, SN = RowTab<E>::SN;
  const double* sg = reinterpret_cast<const double*>(fsm);
  const auto out = gp(out_) + W * lt;
  const int q = int(threadIdx.x), h = q & 1;
#pragma unroll
  for (int v = 0; v < N / 128; ++v) {
    const int k = (q >> 1) + 128 * v;
    *reinterpret_cast<OTM_GLOBAL cd*>(out + long(k) * N + 2 * h) = cd{sg[2 * h * SN + k], sg[(2 * h + 1) * SN + k]};
  }
}

// Pass 1 (row form): a row of R' -> DCT-II along j, written transposed.
template <int E, int W>
__global__ __launch_bounds__(64 * W) void k_fftw_rows_fwd(const Job* __restrict__ jobs,
                                                          const int* __restrict__ starts, int count, bool gate,
                                                          bool rhs) {
  constexpr int N = RowTab<E>::N, M = RowTab<E>::M;
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const RowWave<E, W> G(fsm, L.lt);
  const int t = G.t, row = G.row;
  cd tv[reg_tpt(RowTab<E>::SIZE, W)];
  tab_load<RowTab<E>::SIZE, W>(tv, J.fft);
  cd x[E];
  row_load<E>(x, J.f.r, row, t);
  if (rhs && row == 0 && t == 0) x[0].x -= st->sum_r;  // R' = R - (sum r) e0 e0^T
  if (gate && early && st->done) return;
  tab_store<RowTab<E>::SIZE, W>(tv, G.tab);
  __syncthreads();
  reg_fft<E, -1>(x, G.xb, G.tab, t);
#pragma unroll
  for (int u = 0; u < E; ++u) G.xb[xpad(t + 64 * u)] = x[u];
  __syncthreads();
  const double s0 = sqrt(1.0 / N), s1 = sqrt(2.0 / N);
  double xlo[E], xhi[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int k = t + 64 * u;
    row_post(x[u], G.xb[xpad(k == 0 ? 0 : M - k)], row_w<E>(G.tab, k), row_phase_lo<E>(G.tab, k),
             row_phase_hi<E>(G.tab, k), k == 0 ? s0 : s1, s1, xlo[u], xhi[u]);
  }
  __syncthreads();  // every partner read is done: the exchange buffers become the staging rows
  double* const sg = G.stage(fsm);
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int k = t + 64 * u;
    sg[k] = xlo[u];
    sg[k + M] = xhi[u];
  }
  __syncthreads();
  row_store_transposed<E, W>(fsm, J.f.tmp0, L.lt);
}

// Pass 2 (row form): a row of B^T, DCT-II along i, / lambda, DCT-III along
// i, written transposed back.
template <int E, int W>
__global__ __launch_bounds__(64 * W) void k_fftw_cols(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                      int count, bool gate) {
  constexpr int N = RowTab<E>::N, M = RowTab<E>::M;
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const RowWave<E, W> G(fsm, L.lt);
  const int t = G.t, l = G.row;
  cd tv[reg_tpt(RowTab<E>::SIZE, W)];
  tab_load<RowTab<E>::SIZE, W>(tv, J.fft);
  cd x[E];
  row_load<E>(x, J.f.tmp0, l, t);
  // eigenvalue halves sin^2(pi |k-l| / 2n) + sin^2(pi (k+l) / 2n) of k and k + n/2
  const auto sq = gp(J.sq);
  double lam[2][E];
#pragma unroll
  for (int u = 0; u < E; ++u)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = t + 64 * u + h * M;
      const int dm = k > l ? k - l : l - k;
      lam[h][u] = sq[dm] + sq[k + l];
    }
  if (gate && early && st->done) return;
  tab_store<RowTab<E>::SIZE, W>(tv, G.tab);
  __syncthreads();
  reg_fft<E, -1>(x, G.xb, G.tab, t);
#pragma unroll
  for (int u = 0; u < E; ++u) G.xb[xpad(t + 64 * u)] = x[u];
  __syncthreads();
  const double s0 = sqrt(1.0 / N), s1 = sqrt(2.0 / N);
  const double r0 = sqrt(double(N)), r1 = sqrt(0.5 * N);
  // spectrum / lambda, scaled by 1 / s_k for the inverse: (U[k], U[k + n/2])
  cd w[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int k = t + 64 * u;
    double xlo, xhi;
    row_post(x[u], G.xb[xpad(k == 0 ? 0 : M - k)], row_w<E>(G.tab, k), row_phase_lo<E>(G.tab, k),
             row_phase_hi<E>(G.tab, k), k == 0 ? s0 : s1, s1, xlo, xhi);
    const double ulo = (k == 0 && l == 0) ? 0. : xlo / (2. * lam[0][u]);
    const double uhi = xhi / (2. * lam[1][u]);
    w[u] = cd{ulo * (k == 0 ? r0 : r1), uhi * r1};
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < E; ++u) G.xb[xpad(t + 64 * u)] = w[u];
  __syncthreads();
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int k = t + 64 * u;
    const cd q = G.xb[xpad(k == 0 ? 0 : M - k)];  // U[n/2 - k], U[n - k]
    const double umk = k == 0 ? w[u].y : q.x, unk = k == 0 ? 0. : q.y;
    x[u] = row_pre(w[u].x, w[u].y, umk, unk, row_w<E>(G.tab, k), row_phase_lo<E>(G.tab, k), row_phase_hi<E>(G.tab, k));
  }
  __syncthreads();
  reg_fft<E, 1>(x, G.xb, G.tab, t);  // ends reading its exchange buffer before a barrier
  const double inv = 1.0 / M;
  double* const sg = G.stage(fsm);
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int m = t + 64 * u;
    sg[row_c0<E>(u, m)] = x[u].x * inv;
    sg[row_c1<E>(u, m)] = x[u].y * inv;
  }
  __syncthreads();
  row_store_transposed<E, W>(fsm, J.f.tmp1, L.lt);
}

// Pass 3 (row form): DCT-III along j = d_hat, and the row's |d_hat|^2 partial.
template <int E, int W>
__global__ __launch_bounds__(64 * W) void k_fftw_rows_inv(const Job* __restrict__ jobs,
                                                          const int* __restrict__ starts, int count, bool gate) {
  constexpr int N = RowTab<E>::N, M = RowTab<E>::M;
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const RowWave<E, W> G(fsm, L.lt);
  const int t = G.t, row = G.row;
  cd tv[reg_tpt(RowTab<E>::SIZE, W)];
  tab_load<RowTab<E>::SIZE, W>(tv, J.fft);
  // X[k], X[k + n/2] of the row; X[n/2 - k], X[n - k] come from the lane of
  // n/2 - k through the exchange buffer (k = 0: X[n/2] and 0)
  const auto in = gp((const double*)J.f.tmp1) + long(row) * N;
  double xl[E], xh[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int k = t + 64 * u;
    xl[u] = in[k];
    xh[u] = in[k + M];
  }
  if (gate && early && st->done) return;
  tab_store<RowTab<E>::SIZE, W>(tv, G.tab);
#pragma unroll
  for (int u = 0; u < E; ++u) G.xb[xpad(t + 64 * u)] = cd{xl[u], xh[u]};
  __syncthreads();
  const double r0 = sqrt(double(N)), r1 = sqrt(0.5 * N);
  cd x[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int k = t + 64 * u;
    const cd q = G.xb[xpad(k == 0 ? 0 : M - k)];
    const double xm = k == 0 ? xh[u] : q.x, xn = k == 0 ? 0. : q.y;
    x[u] = row_pre(xl[u] * (k == 0 ? r0 : r1), xh[u] * r1, xm * r1, xn * r1, row_w<E>(G.tab, k),
                   row_phase_lo<E>(G.tab, k), row_phase_hi<E>(G.tab, k));
  }
  __syncthreads();  // partner reads done before the FFT's first exchange
  reg_fft<E, 1>(x, G.xb, G.tab, t);  // ends reading its exchange buffer before a barrier
  // the row through the wave's exchange buffer, written in 16-byte runs
  double* const sg = reinterpret_cast<double*>(G.xb);
  const double inv = 1.0 / M;
  double nrm[1] = {0.};
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int m = t + 64 * u;
    const double a = x[u].x * inv, b = x[u].y * inv;
    sg[row_c0<E>(u, m)] = a;
    sg[row_c1<E>(u, m)] = b;
    nrm[0] += a * a;
    nrm[0] += b * b;
  }
  __syncthreads();
  const auto out = gp(J.f.dhat) + long(row) * N;
#pragma unroll
  for (int v = 0; v < E; ++v) {
    const int c = 2 * (t + 64 * v);
    *reinterpret_cast<OTM_GLOBAL cd*>(out + c) = cd{sg[c], sg[c + 1]};
  }
  wave_sum<1>(nrm);
  if (t == 63) gp(J.f.site(SITE_DNORM))[row] = nrm[0];
}

template <int E, int W>
void launch_row(const JobSet& js, int kind, int pass, bool gate, hipStream_t s) {
  const size_t lds = RowTab<E>::lds_bytes(W);
  static std::atomic<unsigned long long> attr{0};
  int dev = 0;
  OTM_HIP(hipGetDevice(&dev));
  const unsigned long long bit = 1ull << (dev & 63);
  if (!(attr.load() & bit)) {
    for (const void* k : {reinterpret_cast<const void*>(&k_fftw_rows_fwd<E, W>),
                          reinterpret_cast<const void*>(&k_fftw_cols<E, W>),
                          reinterpret_cast<const void*>(&k_fftw_rows_inv<E, W>)})
      OTM_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    attr.fetch_or(bit);
  }
  const dim3 grid(unsigned(js.total[kind])), block(64 * W);
  const int* st = js.starts_of(kind);
  const Job* jobs = js.jobs;
  switch (pass) {
    case 0: hipLaunchKernelGGL((k_fftw_rows_fwd<E, W>), grid, block, lds, s, jobs, st, js.count, gate, true); break;
    case 1: hipLaunchKernelGGL((k_fftw_cols<E, W>), grid, block, lds, s, jobs, st, js.count, gate); break;
    default: hipLaunchKernelGGL((k_fftw_rows_inv<E, W>), grid, block, lds, s, jobs, st, js.count, gate);
  }
  OTM_HIP(hipGetLastError());
}

template <int E>
void launch_row_kind(const JobSet& js, int kind, int pass, bool gate, hipStream_t s) {
  if (js.total[kind] == 0) return;
  launch_row<E, FFT_ROW_W>(js, kind, pass, gate, s);
}

// ---------------------------------------------------------------------------
// Packed workgroup form (n <= 1024 without a register form, launch kind
// KIND_FFT): 256 threads; a sequence of n points (two rows) takes tps = n/4
// threads, each with four frequencies / columns k = tt + u tps, so one
// workgroup holds nseq = fft_pack_seqs(n) = 1024 / n sequences (2 nseq rows:
// 8 at n = 128, 2 at n = 384, 1 from n = 513 up) instead of leaving the
// threads beyond n/4 idle.  LDS: buffers A and B of nseq sequences, then the
// twiddles: (2 nseq + 1) n complex.  Each thread runs the butterflies of its
// own sequence (tt, tt + tps, ...); the stages of all sequences share the
// workgroup barriers.
// ---------------------------------------------------------------------------
struct PackLoc {
  int N, tps, nseq, q, tt, r0;  // r0: first row of the thread's sequence
  bool act;                     // the thread belongs to a sequence inside the grid
};
__device__ inline PackLoc pack_locate(int N, int lt) {
  PackLoc P;
  P.N = N;
  P.tps = N / 4;
  P.nseq = fft_pack_seqs(N);
  P.q = int(threadIdx.x) / P.tps;
  P.tt = int(threadIdx.x) - P.q * P.tps;
  P.r0 = 2 * (P.nseq * lt + P.q);
  P.act = P.q < P.nseq && P.r0 < N;
  return P;
}
// the n stage twiddles into LDS (issued with the other loads, stored later)
constexpr int PACK_TW = FFT_SPLIT_N / 256;  // twiddles per thread at most
__device__ inline void pack_load_tw(cd (&tv)[PACK_TW], const double* table, int N) {
  const auto tw = gp(reinterpret_cast<const cd*>(table));
#pragma unroll
  for (int u = 0; u < PACK_TW; ++u) {
    const int k = threadIdx.x + u * 256;
    tv[u] = k < N ? tw[k] : cd{0., 0.};
  }
}
__device__ inline void pack_store_tw(const cd (&tv)[PACK_TW], cd* twl, int N) {
#pragma unroll
  for (int u = 0; u < PACK_TW; ++u) {
    const int k = threadIdx.x + u * 256;
    if (k < N) twl[k] = tv[u];
  }
}
// rows r0, r0 + 1 at columns tt + u tps (0 outside the grid)
__device__ inline void pack_load_rows(double (&v)[2][KPT], const double* in_, const PackLoc& P) {
  const auto in = gp(in_);
#pragma unroll
  for (int rr = 0; rr < 2; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) v[rr][u] = P.act ? in[long(P.r0 + rr) * P.N + P.tt + u * P.tps] : 0.;
}
__device__ inline void pack_store_permuted(const double (&v)[2][KPT], cd* A, const PackLoc& P) {
  if (!P.act) return;
  cd* z = A + P.q * P.N;
#pragma unroll
  for (int rr = 0; rr < 2; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) reinterpret_cast<double*>(z + perm_pos(P.tt + u * P.tps, P.N))[rr] = v[rr][u];
}

// Pass 1 (packed form): rows of R' -> DCT-II along j, written transposed.
__global__ __launch_bounds__(256) void k_fftp_rows_fwd(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                       int count, bool gate, bool rhs) {
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const PackLoc P = pack_locate(J.f.n, L.lt);
  const int N = P.N;
  cd* A = fsm;
  cd* B = fsm + P.nseq * N;
  cd* twl = fsm + 2 * P.nseq * N;
  cd tv[PACK_TW];
  pack_load_tw(tv, J.fft, N);
  const auto tg = gp(reinterpret_cast<const cd*>(J.fft));
  cd ph[KPT];
#pragma unroll
  for (int u = 0; u < KPT; ++u) ph[u] = P.act ? tg[N + P.tt + u * P.tps] : cd{0., 0.};
  double v[2][KPT];
  pack_load_rows(v, J.f.r, P);
  if (rhs && P.act && P.r0 == 0 && P.tt == 0) v[0][0] -= st->sum_r;  // R' = R - (sum r) e0 e0^T
  if (gate && early && st->done) return;
  pack_store_tw(tv, twl, N);
  pack_store_permuted(v, A, P);
  __syncthreads();
  const cd* z = fft_seq_p<-1>(A, B, twl, N, P.q * N, P.tt, P.tps, P.act);
  if (!P.act) return;
  const Scales sc(N);
  const auto out = gp(J.f.tmp0);
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int k = P.tt + u * P.tps;
    double x0, x1;
    dct2_post(z + P.q * N, ph[u], sc, N, k, x0, x1);
    *reinterpret_cast<OTM_GLOBAL cd*>(out + long(k) * N + P.r0) = cd{x0, x1};
  }
}

// Pass 2 (packed form): DCT-II along i, / lambda, DCT-III along i, written
// transposed back.
__global__ __launch_bounds__(256) void k_fftp_cols(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                   int count, bool gate) {
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const PackLoc P = pack_locate(J.f.n, L.lt);
  const int N = P.N, l0 = P.r0;
  cd* A = fsm;
  cd* B = fsm + P.nseq * N;
  cd* twl = fsm + 2 * P.nseq * N;
  cd tv[PACK_TW];
  pack_load_tw(tv, J.fft, N);
  const auto tg = gp(reinterpret_cast<const cd*>(J.fft));
  cd ph[KPT];
#pragma unroll
  for (int u = 0; u < KPT; ++u) ph[u] = P.act ? tg[N + P.tt + u * P.tps] : cd{0., 0.};
  double v[2][KPT];
  pack_load_rows(v, J.f.tmp0, P);
  const auto sq = gp(J.sq);
  double lam[2][KPT];
#pragma unroll
  for (int rr = 0; rr < 2; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = P.tt + u * P.tps, l = l0 + rr;
      const int dm = k > l ? k - l : l - k;
      lam[rr][u] = P.act ? sq[dm] + sq[k + l] : 1.;
    }
  if (gate && early && st->done) return;
  pack_store_tw(tv, twl, N);
  pack_store_permuted(v, A, P);
  __syncthreads();
  cd* z = fft_seq_p<-1>(A, B, twl, N, P.q * N, P.tt, P.tps, P.act);
  cd* other = z == A ? B : A;
  const Scales sc(N);
  // spectrum / lambda of the sequence's two rows, in its region of the free buffer
  double* X = reinterpret_cast<double*>(other + P.q * N);
  if (P.act) {
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = P.tt + u * P.tps;
      double x[2];
      dct2_post(z + P.q * N, ph[u], sc, N, k, x[0], x[1]);
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) X[rr * N + k] = (k == 0 && l0 + rr == 0) ? 0. : x[rr] / (2. * lam[rr][u]);
    }
  }
  __syncthreads();
  if (P.act) {
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = P.tt + u * P.tps;
      const int kn = k == 0 ? 0 : N - k;
      z[P.q * N + k] = dct3_pre(X[k], X[kn], X[N + k], X[N + kn], ph[u], sc, k);
    }
  }
  __syncthreads();
  const cd* y = fft_seq_p<1>(z, other, twl, N, P.q * N, P.tt, P.tps, P.act);
  if (!P.act) return;
  const auto out = gp(J.f.tmp1);
  const double inv = 1.0 / N;
#pragma unroll
  for (int u = 0; u < KPT; ++u) {
    const int i = P.tt + u * P.tps;
    const cd a = y[P.q * N + perm_pos(i, N)];
    *reinterpret_cast<OTM_GLOBAL cd*>(out + long(i) * N + l0) = cd{a.x * inv, a.y * inv};
  }
}

// Pass 3 (packed form): DCT-III along j = d_hat, and one |d_hat|^2 partial
// per workgroup.
__global__ __launch_bounds__(256) void k_fftp_rows_inv(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                                       int count, bool gate) {
  extern __shared__ cd fsm[];
  const FftLoc L = fft_locate(starts, count);
  const Job& J = jobs[L.p];
  const auto st = gp((const SolveState*)J.f.state);
  const bool early = count == 1;
  if (gate && !early && st->done) return;
  const PackLoc P = pack_locate(J.f.n, L.lt);
  const int N = P.N;
  cd* A = fsm;
  cd* B = fsm + P.nseq * N;
  cd* twl = fsm + 2 * P.nseq * N;
  cd tv[PACK_TW];
  pack_load_tw(tv, J.fft, N);
  const auto tg = gp(reinterpret_cast<const cd*>(J.fft));
  cd ph[KPT];
#pragma unroll
  for (int u = 0; u < KPT; ++u) ph[u] = P.act ? tg[N + P.tt + u * P.tps] : cd{0., 0.};
  // X[row][k] and X[row][N-k] of the sequence's two rows
  const auto in = gp((const double*)J.f.tmp1) + long(P.act ? P.r0 : 0) * N;
  double xk[2][KPT], xn[2][KPT];
#pragma unroll
  for (int rr = 0; rr < 2; ++rr)
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = P.tt + u * P.tps;
      const int kn = k == 0 ? 0 : N - k;
      xk[rr][u] = P.act ? in[rr * N + k] : 0.;
      xn[rr][u] = P.act ? in[rr * N + kn] : 0.;
    }
  if (gate && early && st->done) return;
  pack_store_tw(tv, twl, N);
  const Scales sc(N);
  if (P.act) {
#pragma unroll
    for (int u = 0; u < KPT; ++u) {
      const int k = P.tt + u * P.tps;
      A[P.q * N + k] = dct3_pre(xk[0][u], xn[0][u], xk[1][u], xn[1][u], ph[u], sc, k);
    }
  }
  __syncthreads();
  const cd* y = fft_seq_p<1>(A, B, twl, N, P.q * N, P.tt, P.tps, P.act);
  const double inv = 1.0 / N;
  double nrm[1] = {0.};
  if (P.act) {
    const auto out = gp(J.f.dhat) + long(P.r0) * N;
#pragma unroll
    for (int rr = 0; rr < 2; ++rr)
#pragma unroll
      for (int u = 0; u < KPT; ++u) {
        const int c = P.tt + u * P.tps;
        const cd a = y[P.q * N + perm_pos(c, N)];
        const double x = (rr ? a.y : a.x) * inv;
        out[long(rr) * N + c] = x;
        nrm[0] += x * x;
      }
  }
  // workgroup partial: DPP wave sums, then the wave totals in a fixed order
  // through LDS, in the buffer that does not hold y (the last FFT stage read
  // it before its closing barrier)
  wave_sum<1>(nrm);
  double* red = reinterpret_cast<double*>(y == A ? B : A);
  if ((threadIdx.x & 63) == 63) red[threadIdx.x >> 6] = nrm[0];
  __syncthreads();
  if (threadIdx.x == 0) gp(J.f.site(SITE_DNORM))[L.lt] = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_pack_kind(const JobSet& js, int kind, int pass, bool gate, hipStream_t s) {
  if (js.total[kind] == 0) return;
  size_t lds = 0;
  for (int p = 0; p < js.count; ++p) {
    const int n = js.host[p].f.n;
    if (js.jroute[p] == 2 && fft_pack(n)) lds = std::max(lds, size_t(2 * fft_pack_seqs(n) + 1) * n * sizeof(cd));
  }
  // the kernels use no static LDS; the dynamic size is allowed once per
  // device (bit d of the mask) up to the largest it can take, n = 1024
  static std::atomic<unsigned long long> attr{0};
  int dev = 0;
  OTM_HIP(hipGetDevice(&dev));
  const unsigned long long bit = 1ull << (dev & 63);
  if (!(attr.load() & bit)) {
    const int cap = int(size_t(2 * fft_pack_seqs(FFT_SPLIT_N) + 1) * FFT_SPLIT_N * sizeof(cd));
    for (const void* k : {reinterpret_cast<const void*>(&k_fftp_rows_fwd), reinterpret_cast<const void*>(&k_fftp_cols),
                          reinterpret_cast<const void*>(&k_fftp_rows_inv)})
      OTM_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, cap));
    attr.fetch_or(bit);
  }
  const dim3 grid(unsigned(js.total[kind])), block(256);
  const int* st = js.starts_of(kind);
  const Job* jobs = js.jobs;
  switch (pass) {
    case 0: hipLaunchKernelGGL(k_fftp_rows_fwd, grid, block, lds, s, jobs, st, js.count, gate, true); break;
    case 1: hipLaunchKernelGGL(k_fftp_cols, grid, block, lds, s, jobs, st, js.count, gate); break;
    default: hipLaunchKernelGGL(k_fftp_rows_inv, grid, block, lds, s, jobs, st, js.count, gate);
  }
  OTM_HIP(hipGetLastError());
}

template <int FS>
void launch_pass_kind(const JobSet& js, int kind, int pass, bool gate, hipStream_t s) {
  if (js.total[kind] == 0) return;
  int nmax = 0;
  for (int p = 0; p < js.count; ++p)
    if (js.jroute[p] == 2 && fft_rows(js.host[p].f.n) == 2 * FS && !fft_row_form(js.host[p].f.n))
      nmax = std::max(nmax, js.host[p].f.n);
  const size_t lds = size_t(fft_lds(FS)) * nmax;
  // the kernels use no static LDS: up to fft_lds(FS) times the largest n of
  // the kind dynamic, allowed once per device (bit d of the mask)
  static std::atomic<unsigned long long> attr{0};
  int dev = 0;
  OTM_HIP(hipGetDevice(&dev));
  const unsigned long long bit = 1ull << (dev & 63);
  if (!(attr.load() & bit)) {
    const int cap = fft_lds(FS) * (FS == 1 ? FFT_SPLIT_N : FFT_MAX_N);
    for (const void* k : {reinterpret_cast<const void*>(&k_fft_rows_fwd<FS>),
                          reinterpret_cast<const void*>(&k_fft_cols<FS>),
                          reinterpret_cast<const void*>(&k_fft_rows_inv<FS>)})
      OTM_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, cap));
    attr.fetch_or(bit);
  }
  const dim3 grid(unsigned(js.total[kind])), block(256 * FS);
  const int* st = js.starts_of(kind);
  const Job* jobs = js.jobs;
  switch (pass) {
    case 0: hipLaunchKernelGGL(k_fft_rows_fwd<FS>, grid, block, lds, s, jobs, st, js.count, gate, true); break;
    case 1: hipLaunchKernelGGL(k_fft_cols<FS>, grid, block, lds, s, jobs, st, js.count, gate); break;
    default: hipLaunchKernelGGL(k_fft_rows_inv<FS>, grid, block, lds, s, jobs, st, js.count, gate);
  }
  OTM_HIP(hipGetLastError());
}

}  // namespace

// d_hat of every job on the FFT route: launch kind KIND_FFT (n <= FFT_SPLIT_N,
// one workgroup per 2 rows) and KIND_FFT2 (one per 4 rows); dynamic LDS
// sized for the largest grid of each kind.
void launch_fft_pass(const JobSet& js, int pass, bool gate, hipStream_t s) {
  launch_row_kind<4>(js, KIND_FFTW4, pass, gate, s);
  launch_row_kind<8>(js, KIND_FFTW8, pass, gate, s);
  launch_row_kind<16>(js, KIND_FFTW16, pass, gate, s);
  launch_pack_kind(js, KIND_FFT, pass, gate, s);
  launch_pass_kind<2>(js, KIND_FFT2, pass, gate, s);
}

void launch_fft_solve(const JobSet& js, bool gate, hipStream_t s) {
  for (int pass = 0; pass < 3; ++pass) launch_fft_pass(js, pass, gate, s);
}

}  // namespace otm
