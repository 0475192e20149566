// Mixed-precision Laplacian solve: the same 2-D DCT-II diagonalisation as the
// FP64 path (d_hat = C^T [(C R' C^T) / lambda] C, R' = r - (sum r) e0), with
// every GEMM run as a 3-term BF16 split on MFMA (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation), followed by FP64 iterative refinement against the exact
// operator (weight -1/2 at the four clamped diagonal neighbours, 2 on the
// diagonal; otlib/otsolver_2dgrid.cpp:284-306).
//
// Operands live in HBM already split: a value x is stored as three BF16 planes
// x = x0 + x1 + x2 (x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)),
// so the GEMM inner loop only moves bytes and issues MFMAs.  Every plane of R
// rows by K columns (k contiguous in the products) is stored in 16 x 32
// blocks of 1 KB in MFMA fragment order (bofs below): a wave moves one block
// as one contiguous run, from HBM and through LDS.  A product keeps
// the six terms a_i b_j with i + j <= 2.  The constant DCT factors are split
// once per grid size; every stage splits its own output once in its epilogue.
//
// The transform is even/odd folded (h = n/2; P = [E; O], the even and odd rows
// of C restricted to the first h columns, Pt = P^T), each stage being
// D[m][c] = sum_k A[m][k] B[c][k] with k contiguous in both operands:
//   prep  U = R'(:, j) + R'(:, n-1-j),  W = R'(:, j) - R'(:, n-1-j)     (n x h, FP64 -> planes)
//   M1    T[i][l] = sum_j {U|W}[i][j] P[l][j]            (U for l < h, W for l >= h)
//         -> U2[l][i] = T[i][l] + T[n-1-i][l],  W2 = difference        (rows i and n-1-i in one tile)
//   M2    Z[l][k] = sum_i {U2|W2}[l][i] P[k][i] / lambda  (U2 for k < h)   Z = Y^T, even/odd order
//   M3    T3[i][l] = sum_k Z[l][k] Pt[i][k]: the halves k < h, k >= h give rows i and n-1-i
//   M4    X[i][j]  = sum_l T3[i][l] Pt[j][l]: the halves give columns j and n-1-j (FP64 out)
// Each stage is n^3 multiply-adds (x6 BF16 products) instead of 2 n^3.
//
// Every kernel runs over a job table (one launch for all grids of a batch).

#include <cfloat>
#include <cmath>
#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "otm_internal.h"
#include "reduce.h"
#include "split.h"

namespace otm {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int MTM = 64;        // A rows per workgroup tile
constexpr int MKB = 32;        // K step

enum MStage { M1 = 0, M2 = 1, M3 = 2, M4 = 3 };
// M4 epilogue: x0 = D | d_hat = D (+partials) | d_hat = x0 + D (+partials) | x0 += D
enum MFinal { FIN_X0 = 0, FIN_OUT = 1, FIN_ADD_OUT = 2, FIN_ADD_X0 = 3 };



// One 16-byte global -> LDS transfer per lane (global_load_lds_dwordx4): lane
// l's 16 bytes at sbase + voff land at LDS byte lds + 16 l (lds wave-uniform).
// Issued from asm so that the compiler's wait analysis, which cannot tell the
// K-step buffers apart, does not drain every transfer before each LDS read;
// the kernels wait for them explicitly (vmcnt) before the barrier that
// publishes a buffer.  The instruction has no register results.
__device__ inline void lds_dma16(const void* sbase, unsigned voff, unsigned lds) {
  asm volatile("s_mov_b32 m0, %2\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(sbase), "s"(lds) : "memory", "m0");
}
// Wait until at most N of this wave's vector memory operations are
// outstanding and its LDS reads have completed, then the workgroup barrier
// (also a compiler memory barrier: no memory access moves across it).
// Callers pass the number of transfers this wave issued for the K step that
// may stay in flight.  The count cannot be too permissive: vector memory
// operations of a wave complete in issue order, so the outstanding ones are
// always the most recently issued; the transfers are volatile asm and keep
// their order, and any load or store the compiler places after them only
// makes the wait stricter for the older step.  The build rejects a mixed
// stage kernel that spills to scratch.
template <int N>
__device__ inline void wait_vm_barrier() {
  static_assert(N == 0 || N == 4 || N == 5 || N == 6 || N == 12, "waits used by the mixed stages");
  if (N == 0)
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
  else if (N == 4)
    asm volatile("s_waitcnt vmcnt(4) lgkmcnt(0)\n\ts_barrier" ::: "memory");
  else if (N == 5)
    asm volatile("s_waitcnt vmcnt(5) lgkmcnt(0)\n\ts_barrier" ::: "memory");
  else if (N == 6)
    asm volatile("s_waitcnt vmcnt(6) lgkmcnt(0)\n\ts_barrier" ::: "memory");
  else
    asm volatile("s_waitcnt vmcnt(12) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}


// ---------------------------------------------------------------------------
// right-hand side: fold the columns of r' = r - (sum r) e0 and split them
// into the U / W planes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_mprep(const Job* __restrict__ jobs, const int* __restrict__ starts, int count,
                                             bool gate) {
  const int p = count == 1 ? 0 : job_of(starts, count, blockIdx.x);
  const int lt = count == 1 ? int(blockIdx.x) : blockIdx.x - starts[p];
  const int nb = count == 1 ? int(gridDim.x) : starts[p + 1] - starts[p];
  const Job& J = jobs[p];
  const auto st = gp((const SolveState*)J.f.state);
  if (gate && st->done) return;
  const int n = J.f.n, h = n / 2;
  const long nh = long(n) * h;
  const auto r = gp((const double*)J.f.r);
  const auto out = gp(J.f.mrhs);
  const double S = st->sum_r;
  auto src = [&](int i, int j) {
    const long f = long(i) * n + j;
    double v = r[f];
    if (f == 0) v -= S;
    return v;
  };
  // four consecutive j per thread (h % 4 == 0): 8-byte stores per plane
  for (long q4 = long(lt) * NT + threadIdx.x; q4 < nh / 4; q4 += long(nb) * NT) {
    const long q = q4 * 4;
    const int i = int(q / h), j = int(q - long(i) * h);
    u16 u[3][4], w[3][4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double a = src(i, j + e), b = src(i, n - 1 - j - e);
      split3d(a + b, u[0][e], u[1][e], u[2][e]);
      split3d(a - b, w[0][e], w[1][e], w[2][e]);
    }
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      u32x2 pu, pw;
      pu.x = u[pl][0] | (unsigned(u[pl][1]) << 16); pu.y = u[pl][2] | (unsigned(u[pl][3]) << 16);
      pw.x = w[pl][0] | (unsigned(w[pl][1]) << 16); pw.y = w[pl][2] | (unsigned(w[pl][3]) << 16);
      *reinterpret_cast<OTM_GLOBAL u32x2*>(out + pl * nh + bofs(i, j, h)) = pu;
      *reinterpret_cast<OTM_GLOBAL u32x2*>(out + (3 + pl) * nh + bofs(i, j, h)) = pw;
    }
  }
}

// ---------------------------------------------------------------------------
// refinement right-hand side: s = r' - M0 x0 with the exact operator (2 on
// the diagonal, -1/2 at the four clamped diagonal neighbours), folded over
// columns (U = s(:, j) + s(:, n-1-j), W = difference) and split into the six
// BF16 planes.  One workgroup per RT rows x RT folded columns: x0 of the rows
// i0-1 .. i0+RT and of the columns j0-1 .. j0+RT and their mirrors is staged
// in LDS with clamped indices; each thread folds four consecutive columns of
// one row.  s(i, j) = (r(i, j) - [i = j = 0] sum r) - (2 x0(i, j) - 0.5
// ((x0(i-1, j-1) + x0(i-1, j+1)) + (x0(i+1, j-1) + x0(i+1, j+1)))).
// ---------------------------------------------------------------------------
constexpr int RT = 32;
static_assert(RT * RT / 4 == NT, "four folded columns per thread");

__global__ __launch_bounds__(NT) void k_mresid(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                              int count, bool gate) {
  const int p = count == 1 ? 0 : job_of(starts, count, blockIdx.x);
  const int lt = count == 1 ? int(blockIdx.x) : blockIdx.x - starts[p];
  const Job& J = jobs[p];
  const auto st = gp((const SolveState*)J.f.state);
  if (gate && st->done) return;
  const int n = J.f.n, h = n / 2;
  const long nh = long(n) * h;
  const int tj = h / RT;
  const int i0 = (lt / tj) * RT, j0 = (lt % tj) * RT;
  const auto r = gp((const double*)J.f.r);
  const auto x = gp((const double*)J.f.x0);
  const auto out = gp(J.f.mrhs);
  constexpr int HW = RT + 2;
  // xl[a][b] = x0(i0-1+a, j0-1+b), xr[a][b] = x0(i0-1+a, n-1-(j0+RT)+b), indices clamped to the grid
  __shared__ double xl[HW][HW + 1], xr[HW][HW + 1];
  const int t = threadIdx.x;
  const int a = t / (RT / 4), c0 = (t % (RT / 4)) * 4;
  const int i = i0 + a;
  // this thread's residuals, loaded before the halo so that both latencies overlap
  // (32-byte aligned runs: j0 + c0 and n - 4 - j0 - c0 are multiples of 4)
  double rl[4], rr[4];
  {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const auto pl = reinterpret_cast<OTM_GLOBAL const d2*>(r + long(i) * n + j0 + c0);
    const auto pr = reinterpret_cast<OTM_GLOBAL const d2*>(r + long(i) * n + (n - 4 - j0 - c0));
    const d2 l0 = pl[0], l1 = pl[1], r0 = pr[0], r1 = pr[1];
    rl[0] = l0.x; rl[1] = l0.y; rl[2] = l1.x; rl[3] = l1.y;
    rr[0] = r1.y; rr[1] = r1.x; rr[2] = r0.y; rr[3] = r0.x;  // columns n-1-j0-c0-e, e = 0..3
  }
  constexpr int HR2 = (HW * HW + NT - 1) / NT;
  double vl[HR2], vr[HR2];
#pragma unroll
  for (int u = 0; u < HR2; ++u) {
    const int e = t + u * NT;
    const int ha = e / HW, hb = e % HW;
    const int gi = min(max(i0 - 1 + ha, 0), n - 1);
    const int gl = min(max(j0 - 1 + hb, 0), n - 1);
    const int gr = min(max(n - 1 - (j0 + RT) + hb, 0), n - 1);
    vl[u] = e < HW * HW ? x[long(gi) * n + gl] : 0.;
    vr[u] = e < HW * HW ? x[long(gi) * n + gr] : 0.;
  }
  const double S = st->sum_r;
#pragma unroll
  for (int u = 0; u < HR2; ++u) {
    const int e = t + u * NT;
    if (e < HW * HW) {
      xl[e / HW][e % HW] = vl[u];
      xr[e / HW][e % HW] = vr[u];
    }
  }
  __syncthreads();
  u16 uu[3][4], ww[3][4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int c = c0 + e, j = j0 + c;
    // left value s(i, j): halo (a + 1, c + 1)
    double va = rl[e];
    if (i == 0 && j == 0) va -= S;
    const double nl = (xl[a][c] + xl[a][c + 2]) + (xl[a + 2][c] + xl[a + 2][c + 2]);
    va = va - (2.0 * xl[a + 1][c + 1] - 0.5 * nl);
    // mirrored value s(i, n-1-j): halo column RT - c
    const int m = RT - c;
    double vb = rr[e];
    const double nr = (xr[a][m - 1] + xr[a][m + 1]) + (xr[a + 2][m - 1] + xr[a + 2][m + 1]);
    vb = vb - (2.0 * xr[a + 1][m] - 0.5 * nr);
    split3d(va + vb, uu[0][e], uu[1][e], uu[2][e]);
    split3d(va - vb, ww[0][e], ww[1][e], ww[2][e]);
  }
  const long q = bofs(i, j0 + c0, h);
#pragma unroll
  for (int pl = 0; pl < 3; ++pl) {
    u32x2 pu, pw;
    pu.x = uu[pl][0] | (unsigned(uu[pl][1]) << 16); pu.y = uu[pl][2] | (unsigned(uu[pl][3]) << 16);
    pw.x = ww[pl][0] | (unsigned(ww[pl][1]) << 16); pw.y = ww[pl][2] | (unsigned(ww[pl][3]) << 16);
    *reinterpret_cast<OTM_GLOBAL u32x2*>(out + pl * nh + q) = pu;
    *reinterpret_cast<OTM_GLOBAL u32x2*>(out + (3 + pl) * nh + q) = pw;
  }
}

// ---------------------------------------------------------------------------
// one stage of the folded transform for every job.  Workgroup tile: 64 rows
// of A (M1: 32 rows and their 32 mirrors) by TN rows of B (TN = 32 or 64,
// picked per stage by the autotuner), K steps of 32.  The workgroup is two
// K groups of four waves (32 x TN/2 each): group 0 accumulates the first
// ceil(S/2) of the S = K/32 steps, group 1 the rest, each through its own
// three LDS buffers filled straight from global memory (global_load_lds, two
// K steps ahead), both in lockstep on one barrier per step.  For M3/M4 the split is
// exactly k < h / k >= h, the two halves the fold needs; for M1/M2 the
// epilogue adds the two FP32 partials.  The order of every sum depends on n
// only, so all tile widths give bit-identical results.
// ---------------------------------------------------------------------------
constexpr int MNT = 2 * NT;  // threads of a mixed GEMM workgroup

template <int ST, int TN>
__global__ __launch_bounds__(MNT) void k_mgemm(const Job* __restrict__ jobs, const int* __restrict__ starts, int count,
                                              bool gate, int fin) {
  const int p = count == 1 ? 0 : job_of(starts, count, blockIdx.x);
  const int lt = count == 1 ? int(blockIdx.x) : blockIdx.x - starts[p];
  const int nbk = count == 1 ? int(gridDim.x) : starts[p + 1] - starts[p];
  const Job& J = jobs[p];
  const auto st = gp((const SolveState*)J.f.state);
  // the gate is tested after the first operand loads are issued (they are in
  // bounds either way), so its latency overlaps theirs
  const int done = gate ? st->done : 0;
  const int n = J.f.n, h = n / 2;
  const long nh = long(n) * h, nn = long(n) * n;
  constexpr int TJ = TN / 32;                   // MFMA tiles per wave along B
  constexpr int BCHUNKS = 3 * TN * 4;           // 16-byte B chunks per K step
  constexpr int BCH = (BCHUNKS + NT - 1) / NT;  // per thread of a group
  const int ntn = (ST == M1 || ST == M2 ? n : h) / TN;
  const int ntm = ST == M1 ? h / (MTM / 2) : n / MTM;
  if (lt >= ntm * ntn) {  // padding workgroup of the job: neutral partial only
    if (ST == M4 && (fin == FIN_OUT || fin == FIN_ADD_OUT)) {
      double z[1] = {0.};
      write_partials_at<1, MNT>(z, J.f.site(SITE_DNORM), lt, nbk);
    }
    return;
  }
  int tm, tn;
  xcd_tile(lt, ntm, ntn, tm, tn);
  const int m0 = tm * (ST == M1 ? MTM / 2 : MTM), n0 = tn * TN;
  // operands (u16 planes): A rows x lda, plane stride aps; B rows x ldb, plane stride bps
  const u16* Ab;
  const u16* Bb;
  int lda, ldb;
  long aps, bps;
  int K;
  if (ST == M1) {
    Ab = J.f.mrhs + (n0 >= h ? 3 * nh : 0); lda = h; aps = nh;
    Bb = J.Pp; ldb = h; bps = nh; K = h;
  } else if (ST == M2) {
    Ab = J.f.mwa + (n0 >= h ? 3 * nh : 0); lda = h; aps = nh;
    Bb = J.Pp; ldb = h; bps = nh; K = h;
  } else if (ST == M3) {
    Ab = J.f.mwb; lda = n; aps = nn;
    Bb = J.Ptp; ldb = n; bps = nh; K = n;
  } else {
    Ab = J.f.mwa; lda = n; aps = nn;
    Bb = J.Ptp; ldb = n; bps = nh; K = n;
  }
  const auto A = gp(Ab);
  const auto B = gp(Bb);

  // LDS per group: three K-step buffers of [A: 3 planes x 4 row blocks][B: 3
  // planes x TN/16 row blocks] blocks of 1 KB (the blocks of the HBM layout);
  // the epilogue reuses the space for the two groups' FP32 tiles [64][TN + 1]
  constexpr int BRB = TN / 16;               // 16-row blocks of B
  constexpr int NBLK = 3 * 4 + 3 * BRB;      // blocks per K step (18 or 24)
  constexpr int DMAW = (NBLK + 3) / 4;       // transfers per wave per K step (at most)
  constexpr int GBUF = NBLK * 1024;          // bytes per buffer
  __shared__ __attribute__((aligned(16))) unsigned char smem[2 * 3 * GBUF];
  static_assert(2 * 3 * GBUF >= 2 * MTM * (TN + 1) * 4, "epilogue staging must fit");
  // group and wave indices are wave-uniform (readfirstlane tells the compiler)
  const int t = threadIdx.x, g = __builtin_amdgcn_readfirstlane(t / NT), tg = t % NT, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(tg >> 6);
  const int wm = (w >> 1) * 32, wn = (w & 1) * (TN / 2);
  const int lr = lane & 15, lq = lane >> 4;  // MFMA output layout: lane (lr, lq) holds rows 4 lq .. + 3 of column lr
  auto arow = [&](int row) {  // global row of tile row `row`
    if (ST == M1) return row < MTM / 2 ? m0 + row : n - m0 - MTM + row;  // rows i and mirrored n-1-i
    return m0 + row;
  };
  // transfer u of this lane: block b = 4u + w of the group's step (b < 12: A
  // plane b / 4, row block b % 4; else B plane (b - 12) / BRB, row block
  // (b - 12) % BRB), at position `lane`: byte offset from the operand base, K block 0
  unsigned voff[DMAW];
#pragma unroll
  for (int u = 0; u < DMAW; ++u) {
    const int b = min(4 * u + w, NBLK - 1);
    long e;
    if (b < 12)
      e = (b >> 2) * aps + bofs(arow(16 * (b & 3)), 0, lda);
    else
      e = ((b - 12) / BRB) * bps + bofs(n0 + 16 * ((b - 12) % BRB), 0, ldb);
    voff[u] = unsigned((e + lane * 8) * 2);
  }
  const unsigned lds_g = unsigned(size_t((__attribute__((address_space(3))) unsigned char*)smem)) + g * 3 * GBUF;
  const unsigned wlds = lds_g + unsigned(w) * 1024;
  // the group's K steps: [kb0, kb0 + mine) in blocks of 32 k
  const int S = K / MKB, S0 = (S + 1) >> 1, mine = g == 0 ? S0 : S - S0;
  const int kb0 = g * S0;
  auto issue = [&](int s, int buf) {
    const u16* a = Ab + (kb0 + s) * 512;
    const u16* bp = Bb + (kb0 + s) * 512;
#pragma unroll
    for (int u = 0; u < DMAW; ++u)
      if (4 * u + w < NBLK) lds_dma16(4 * u < 12 ? (const void*)a : (const void*)bp, voff[u], wlds + buf * GBUF + u * 4096);
  };
  auto frag = [&](int buf, int blk) {
    return *reinterpret_cast<const bf8*>(smem + (g * 3 + buf) * GBUF + ((blk * 64 + lane) << 4));
  };

  auto compute = [&](f4 (&c)[2][TJ], int buf) {
    bf8 fa[2][3], fb[TJ][3];
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
#pragma unroll
      for (int i = 0; i < 2; ++i) fa[i][pl] = frag(buf, pl * 4 + wm / 16 + i);
#pragma unroll
      for (int j = 0; j < TJ; ++j) fb[j][pl] = frag(buf, 12 + pl * BRB + wn / 16 + j);
    }
    // small products first, the leading one last
    constexpr int PA[6] = {2, 1, 0, 1, 0, 0};
    constexpr int PB[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
    for (int q = 0; q < 6; ++q)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          c[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][PA[q]], fb[j][PB[q]], c[i][j], 0, 0, 0);
  };

  // Each group runs S0 bodies in lockstep with the other (the same
  // barriers); body s (buffer s % 3): wait for this wave's transfers of step
  // s (those of step s + 1 may stay in flight) and its LDS reads, barrier
  // (every wave's transfers of step s are in LDS, and every wave is done
  // reading buffer (s + 2) % 3, last read at body s - 1), start the
  // transfers of step s + 2 into that buffer, multiply step s.  When S is odd
  // the last body of group 1 multiplies nothing.
  f4 acc[2][TJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // M2: eigenvalue factors of the tile's rows and columns, staged in LDS for the epilogue
  __shared__ double cfac[MTM + TN];
  double cf = 0.;
  if (ST == M2 && t < MTM + TN) cf = gp(J.ck)[t < MTM ? m0 + t : n0 + t - MTM];
  if (mine > 0) issue(0, 0);
  if (mine > 1) issue(1, 1);
  if (done) {  // no transfer may outlive the workgroup
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    return;
  }
  if (ST == M2 && t < MTM + TN) cfac[t] = cf;
  const int mycount = (NBLK - w + 3) / 4;  // this wave's transfers per step
  auto body = [&](int s, int buf) {
    if (s + 1 >= mine)
      wait_vm_barrier<0>();
    else if (mycount == 6)
      wait_vm_barrier<6>();
    else if (mycount == 5)
      wait_vm_barrier<5>();
    else
      wait_vm_barrier<4>();
    if (s + 2 < mine) issue(s + 2, (buf + 2) % 3);
    if (s < mine) compute(acc, buf);
  };
  int s = 0;
  for (; s + 3 <= S0; s += 3) {
    body(s, 0);
    body(s + 1, 1);
    body(s + 2, 2);
  }
  if (s < S0) body(s, 0);
  if (s + 1 < S0) body(s + 1, 1);
  __syncthreads();

  // stage both groups' FP32 tiles: lane (lr, lq) of MFMA tile (i, j) holds
  // D[wm + 16i + 4 lq + v][wn + 16j + lr]; D0 = group 0 (k < h for M3/M4), D1 = group 1
  constexpr int DL = TN + 1;
  float* const D0 = reinterpret_cast<float*>(smem);
  float* const D1 = D0 + MTM * DL;
  {
    float* Dg = g == 0 ? D0 : D1;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) Dg[(wm + 16 * i + 4 * lq + v) * DL + wn + 16 * j + lr] = acc[i][j][v];
  }
  __syncthreads();

  if (ST == M4) {
    // X[row][col] and X[row][n-1-col] = p0 +- p1 in FP64
    const auto x0 = gp(J.f.x0);
    const auto dh = gp(J.f.dhat);
    double nrm[1] = {0.};
    for (int it = t; it < MTM * TN; it += MNT) {
      const int r = it / TN, c = it % TN;
      const long row = m0 + r;
      const int col = n0 + c;
      const double p0 = double(D0[r * DL + c]), p1 = double(D1[r * DL + c]);
      const long ia = row * n + col, ib = row * n + (n - 1 - col);
      double va = p0 + p1, vb = p0 - p1;
      if (fin == FIN_ADD_OUT || fin == FIN_ADD_X0) {
        va = x0[ia] + va;
        vb = x0[ib] + vb;
      }
      if (fin == FIN_X0 || fin == FIN_ADD_X0) {
        x0[ia] = va;
        x0[ib] = vb;
      } else {
        dh[ia] = va;
        dh[ib] = vb;
        nrm[0] += va * va + vb * vb;
      }
    }
    if (fin == FIN_OUT || fin == FIN_ADD_OUT) write_partials_at<1, MNT>(nrm, J.f.site(SITE_DNORM), lt, nbk);
    return;
  }

  // the other stages write BF16 planes (blocked layout) in 8-16 byte runs
  if (ST == M1) {
    // U2 / W2 [l = n0 + c][i = m0 + a .. a+3]: pair tile row a with its mirror 63 - a
    const auto out = gp(J.f.mwa);
    for (int it = t; it < TN * 8; it += MNT) {
      const int c = it >> 3, a0 = (it & 7) * 4;
      u16 u[3][4], wv[3][4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rt = a0 + q, rm = MTM - 1 - a0 - q;
        const float top = D0[rt * DL + c] + D1[rt * DL + c], mir = D0[rm * DL + c] + D1[rm * DL + c];
        split3(top + mir, u[0][q], u[1][q], u[2][q]);
        split3(top - mir, wv[0][q], wv[1][q], wv[2][q]);
      }
      const long base = bofs(n0 + c, m0 + a0, h);
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        u32x2 pu, pw;
        pu.x = u[pl][0] | (unsigned(u[pl][1]) << 16); pu.y = u[pl][2] | (unsigned(u[pl][3]) << 16);
        pw.x = wv[pl][0] | (unsigned(wv[pl][1]) << 16); pw.y = wv[pl][2] | (unsigned(wv[pl][3]) << 16);
        *reinterpret_cast<OTM_GLOBAL u32x2*>(out + pl * nh + base) = pu;
        *reinterpret_cast<OTM_GLOBAL u32x2*>(out + 3 * nh + pl * nh + base) = pw;
      }
    }
  } else if (ST == M2) {
    // Z[l = m0 + r][k = n0 + c0 .. c0+7] / lambda(perm k, perm l), with
    // lambda = 2 (1 - cos(pi k'/n) cos(pi l'/n)) formed in FP64 (relative
    // error ~1e-16 / (pi^2 / n^2)) and the division in FP32
    const auto out = gp(J.f.mwb);
    for (int it = t; it < MTM * TN / 8; it += MNT) {
      const int r = it / (TN / 8), c0 = (it % (TN / 8)) * 8;
      const int lrow = m0 + r;
      const double cl = cfac[r];
      u16 z[3][8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int kc = n0 + c0 + q;
        const float lam = float(2. * (1. - cfac[MTM + c0 + q] * cl));
        const float y = D0[r * DL + c0 + q] + D1[r * DL + c0 + q];
        const float val = (kc == 0 && lrow == 0) ? 0.f : y / lam;
        split3(val, z[0][q], z[1][q], z[2][q]);
      }
      const long base = bofs(lrow, n0 + c0, n);
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        u32x4 pz;
        pz.x = z[pl][0] | (unsigned(z[pl][1]) << 16); pz.y = z[pl][2] | (unsigned(z[pl][3]) << 16);
        pz.z = z[pl][4] | (unsigned(z[pl][5]) << 16); pz.w = z[pl][6] | (unsigned(z[pl][7]) << 16);
        *reinterpret_cast<OTM_GLOBAL u32x4*>(out + pl * nn + base) = pz;
      }
    }
  } else {
    // T3[i = n0 + c][l = m0 + r0 .. r0+7] = p0 + p1 and T3[n-1-i][l] = p0 - p1
    const auto out = gp(J.f.mwa);
    for (int it = t; it < TN * 8; it += MNT) {
      const int c = it >> 3, r0 = (it & 7) * 8;
      u16 s0[3][8], s1[3][8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float lo = D0[(r0 + q) * DL + c], hi = D1[(r0 + q) * DL + c];
        split3(lo + hi, s0[0][q], s0[1][q], s0[2][q]);
        split3(lo - hi, s1[0][q], s1[1][q], s1[2][q]);
      }
      const long b0 = bofs(n0 + c, m0 + r0, n), b1 = bofs(n - 1 - n0 - c, m0 + r0, n);
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        u32x4 p0, p1;
        p0.x = s0[pl][0] | (unsigned(s0[pl][1]) << 16); p0.y = s0[pl][2] | (unsigned(s0[pl][3]) << 16);
        p0.z = s0[pl][4] | (unsigned(s0[pl][5]) << 16); p0.w = s0[pl][6] | (unsigned(s0[pl][7]) << 16);
        p1.x = s1[pl][0] | (unsigned(s1[pl][1]) << 16); p1.y = s1[pl][2] | (unsigned(s1[pl][3]) << 16);
        p1.z = s1[pl][4] | (unsigned(s1[pl][5]) << 16); p1.w = s1[pl][6] | (unsigned(s1[pl][7]) << 16);
        *reinterpret_cast<OTM_GLOBAL u32x4*>(out + pl * nn + b0) = p0;
        *reinterpret_cast<OTM_GLOBAL u32x4*>(out + pl * nn + b1) = p1;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Wide form of a stage for large launches (configs 2 and 3): 128 x 128
// workgroup tiles (M1: 64 rows and their 64 mirrors), four waves of 64 x 64
// (96 MFMAs per 24 LDS fragment reads) or eight waves of 64 x 32 (two waves
// per SIMD, 48 MFMAs per 18 fragment reads).  Operand tiles go from global
// memory straight into LDS (global_load_lds_dwordx4: no register staging, no
// LDS store instructions), three K-step buffers deep; every transfer and every
// fragment read is one contiguous 1 KB block.  The K steps of the two
// halves (k < K/2 into `lo`, k >= K/2 into `acc`) are processed alternately,
// each half in increasing k, so every output accumulates in the same order as
// the two-group form and the epilogue adds (or subtracts) the two FP32 halves
// exactly as it does: results are bit-identical.
// ---------------------------------------------------------------------------
constexpr int WT = 128;  // wide tile rows / columns


template <int ST, int NW>
__global__ __launch_bounds__(NW * 64) void k_mgemm_w(const Job* __restrict__ jobs, const int* __restrict__ starts,
                                               int count, bool gate, int fin) {
  const int p = count == 1 ? 0 : job_of(starts, count, blockIdx.x);
  const int lt = count == 1 ? int(blockIdx.x) : blockIdx.x - starts[p];
  const int nbk = count == 1 ? int(gridDim.x) : starts[p + 1] - starts[p];
  const Job& J = jobs[p];
  const auto st = gp((const SolveState*)J.f.state);
  const int done = gate ? st->done : 0;
  const int n = J.f.n, h = n / 2;
  const long nh = long(n) * h, nn = long(n) * n;
  constexpr int WTH = NW * 64;                 // threads
  constexpr int TI = 4, TJ = NW == 4 ? 4 : 2;  // MFMA tiles per wave (64 x 64 or 64 x 32)
  const int ntn = (ST == M1 || ST == M2 ? n : h) / WT;
  const int ntm = ST == M1 ? h / (WT / 2) : n / WT;
  if (lt >= ntm * ntn) {
    if (ST == M4 && (fin == FIN_OUT || fin == FIN_ADD_OUT)) {
      double z[1] = {0.};
      write_partials_at<1, WTH>(z, J.f.site(SITE_DNORM), lt, nbk);
    }
    return;
  }
  int tm, tn;
  xcd_tile(lt, ntm, ntn, tm, tn);
  const int m0 = tm * (ST == M1 ? WT / 2 : WT), n0 = tn * WT;
  const u16* Ab;
  const u16* Bb;
  int lda, ldb;
  long aps, bps;
  int K;
  if (ST == M1) {
    Ab = J.f.mrhs + (n0 >= h ? 3 * nh : 0); lda = h; aps = nh;
    Bb = J.Pp; ldb = h; bps = nh; K = h;
  } else if (ST == M2) {
    Ab = J.f.mwa + (n0 >= h ? 3 * nh : 0); lda = h; aps = nh;
    Bb = J.Pp; ldb = h; bps = nh; K = h;
  } else if (ST == M3) {
    Ab = J.f.mwb; lda = n; aps = nn;
    Bb = J.Ptp; ldb = n; bps = nh; K = n;
  } else {
    Ab = J.f.mwa; lda = n; aps = nn;
    Bb = J.Ptp; ldb = n; bps = nh; K = n;
  }
  const auto A = gp(Ab);
  const auto B = gp(Bb);

  // LDS: three K-step buffers, each [operand A, B][plane 3][16-row block 8]
  // [lane 64] x 16 bytes: the blocks of the HBM layout, so a transfer moves
  // one contiguous 1 KB block and a fragment read is one contiguous 1 KB run
  constexpr int SLOTS = 2 * 3 * WT * 4;  // 16-byte chunks per K step
  constexpr int BUF = SLOTS * 16;        // bytes per buffer
  constexpr int DMA = SLOTS / WTH;       // transfers per thread per K step (12 or 6)
  __shared__ __attribute__((aligned(16))) unsigned char smem[3 * BUF];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = NW == 4 ? (w >> 1) * 64 : (w >> 2) * 64, wn = NW == 4 ? (w & 1) * 64 : (w & 3) * 32;
  const int lr = lane & 15, lq = lane >> 4;  // MFMA output layout: lane (lr, lq) holds rows 4 lq .. + 3 of column lr
  auto arow = [&](int row) {
    if (ST == M1) return row < WT / 2 ? m0 + row : n - m0 - WT + row;
    return m0 + row;
  };

  __shared__ double cfac[2 * WT];
  double cf = 0.;
  if (ST == M2 && t < 2 * WT) cf = gp(J.ck)[t < WT ? m0 + t : n0 + t - WT];

  // transfer u of this lane (u < DMA / 2: A, else B) fills LDS block NW u + w
  // of the step (plane ((NW u + w) % 24) / 8, row block (NW u + w) % 8) at
  // position `lane`: its byte offset from the operand base, K step 0
  unsigned voff[DMA];
#pragma unroll
  for (int u = 0; u < DMA; ++u) {
    const int blk = (NW * u + w) % 24, pl = blk >> 3, rb = blk & 7;
    const long e = u < DMA / 2 ? pl * aps + bofs(arow(16 * rb), 0, lda) : pl * bps + bofs(n0 + 16 * rb, 0, ldb);
    voff[u] = unsigned((e + lane * 8) * 2);
  }
  const unsigned lds0 = unsigned(size_t((__attribute__((address_space(3))) unsigned char*)smem));
  const unsigned wlds = lds0 + unsigned(__builtin_amdgcn_readfirstlane(w)) * 1024;
  const int S = K / MKB, S0 = S >> 1;
  // K step of position q in the alternating order: 0, S0, 1, S0 + 1, ...
  auto kstep = [&](int q) { return (q & 1) ? S0 + (q >> 1) : (q >> 1); };
  auto issue = [&](int q, int buf) {
    const int kb = kstep(q);  // K block: 512 elements further along every row block
    const u16* a = Ab + kb * 512;
    const u16* b = Bb + kb * 512;
#pragma unroll
    for (int u = 0; u < DMA; ++u)
      lds_dma16(u < DMA / 2 ? (const void*)a : (const void*)b, voff[u], wlds + buf * BUF + u * NW * 1024);
  };
  // fragment of 16-row block rb of plane pl of operand op
  auto frag = [&](int buf, int op, int pl, int rb) {
    return *reinterpret_cast<const bf8*>(smem + buf * BUF + op * (BUF / 2) + (((pl * 8 + rb) * 64 + lane) << 4));
  };
  f4 acc[TI][TJ], lo[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      lo[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }
  // Position q (buffer q % 3): wait for this wave's transfers of q (issued
  // two positions earlier; those of q + 1 may stay in flight), barrier (every
  // wave's transfers of q are in LDS, and every wave is done reading buffer
  // (q + 2) % 3, last read at position q - 1), start the transfers of q + 2,
  // read the fragments and issue the 96 MFMAs.
  auto body = [&](f4 (&c)[TI][TJ], int q, int buf) {
    if (q + 1 < S)
      wait_vm_barrier<DMA>();
    else
      wait_vm_barrier<0>();
    if (q + 2 < S) issue(q + 2, (buf + 2) % 3);
    bf8 fa[TI][3], fb[TJ][3];
    constexpr int RA[3] = {2, 1, 0};
    constexpr int RB[3] = {0, 1, 2};
#pragma unroll
    for (int x = 0; x < 3; ++x) {
#pragma unroll
      for (int i = 0; i < TI; ++i) fa[i][RA[x]] = frag(buf, 0, RA[x], wm / 16 + i);
#pragma unroll
      for (int j = 0; j < TJ; ++j) fb[j][RB[x]] = frag(buf, 1, RB[x], wn / 16 + j);
    }
    constexpr int PA[6] = {2, 1, 0, 1, 0, 0};
    constexpr int PB[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
    for (int x = 0; x < 6; ++x)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          c[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][PA[x]], fb[j][PB[x]], c[i][j], 0, 0, 0);
  };
  issue(0, 0);
  issue(1, 1);
  if (done) {  // no transfer may outlive the workgroup
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    return;
  }
  if (ST == M2 && t < 2 * WT) cfac[t] = cf;
  // S is a multiple of 4: six positions per round (buffers 0, 1, 2, 0, 1, 2),
  // then 2 or 4 more (buffers 0, 1 and 2, 0)
  int q = 0;
  for (; q + 6 <= S; q += 6) {
    body(lo, q, 0);
    body(acc, q + 1, 1);
    body(lo, q + 2, 2);
    body(acc, q + 3, 0);
    body(lo, q + 4, 1);
    body(acc, q + 5, 2);
  }
  if (q + 2 <= S) {
    body(lo, q, 0);
    body(acc, q + 1, 1);
    q += 2;
  }
  if (q + 2 <= S) {
    body(lo, q, 2);
    body(acc, q + 1, 0);
  }
  __syncthreads();  // every wave is done reading the buffers before the epilogue reuses them

  // lane (lr, lq) of tile (i, j) holds D[wm + 16i + 4lq + v][wn + 16j + lr]; lo = k-steps < S0, acc = the rest
  if (ST == M4) {
    const auto x0 = gp(J.f.x0);
    const auto dh = gp(J.f.dhat);
    double nrm[1] = {0.};
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const long row = m0 + wm + 16 * i + 4 * lq + v;
          const int col = n0 + wn + 16 * j + lr;
          const double p0 = double(lo[i][j][v]), p1 = double(acc[i][j][v]);
          const long ia = row * n + col, ib = row * n + (n - 1 - col);
          double va = p0 + p1, vb = p0 - p1;
          if (fin == FIN_ADD_OUT || fin == FIN_ADD_X0) {
            va = x0[ia] + va;
            vb = x0[ib] + vb;
          }
          if (fin == FIN_X0 || fin == FIN_ADD_X0) {
            x0[ia] = va;
            x0[ib] = vb;
          } else {
            dh[ia] = va;
            dh[ib] = vb;
            nrm[0] += va * va + vb * vb;
          }
        }
    if (fin == FIN_OUT || fin == FIN_ADD_OUT) write_partials_at<1, WTH>(nrm, J.f.site(SITE_DNORM), lt, nbk);
    return;
  }

  // the other stages stage one FP32 tile [128][132] in LDS per round
  constexpr int DL = WT + 4;
  float* const Ds = reinterpret_cast<float*>(smem);
  auto stage = [&](float sign) {
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          Ds[(wm + 16 * i + 4 * lq + v) * DL + wn + 16 * j + lr] =
              sign > 0.f ? lo[i][j][v] + acc[i][j][v] : lo[i][j][v] - acc[i][j][v];
    __syncthreads();
  };
  if (ST == M1) {
    stage(1.f);
    const auto out = gp(J.f.mwa);
    for (int it = t; it < WT * (WT / 8); it += WTH) {
      const int c = it / (WT / 8), a0 = (it % (WT / 8)) * 4;
      u16 u[3][4], wv[3][4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float top = Ds[(a0 + q) * DL + c], mir = Ds[(WT - 1 - a0 - q) * DL + c];
        split3(top + mir, u[0][q], u[1][q], u[2][q]);
        split3(top - mir, wv[0][q], wv[1][q], wv[2][q]);
      }
      const long base = bofs(n0 + c, m0 + a0, h);
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        u32x2 pu, pw;
        pu.x = u[pl][0] | (unsigned(u[pl][1]) << 16); pu.y = u[pl][2] | (unsigned(u[pl][3]) << 16);
        pw.x = wv[pl][0] | (unsigned(wv[pl][1]) << 16); pw.y = wv[pl][2] | (unsigned(wv[pl][3]) << 16);
        *reinterpret_cast<OTM_GLOBAL u32x2*>(out + pl * nh + base) = pu;
        *reinterpret_cast<OTM_GLOBAL u32x2*>(out + 3 * nh + pl * nh + base) = pw;
      }
    }
  } else if (ST == M2) {
    stage(1.f);
    const auto out = gp(J.f.mwb);
    for (int it = t; it < WT * WT / 8; it += WTH) {
      const int r = it / (WT / 8), c0 = (it % (WT / 8)) * 8;
      const int lrow = m0 + r;
      const double cl = cfac[r];
      u16 z[3][8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int kc = n0 + c0 + q;
        const float lam = float(2. * (1. - cfac[WT + c0 + q] * cl));
        const float y = Ds[r * DL + c0 + q];
        const float val = (kc == 0 && lrow == 0) ? 0.f : y / lam;
        split3(val, z[0][q], z[1][q], z[2][q]);
      }
      const long base = bofs(lrow, n0 + c0, n);
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {
        u32x4 pz;
        pz.x = z[pl][0] | (unsigned(z[pl][1]) << 16); pz.y = z[pl][2] | (unsigned(z[pl][3]) << 16);
        pz.z = z[pl][4] | (unsigned(z[pl][5]) << 16); pz.w = z[pl][6] | (unsigned(z[pl][7]) << 16);
        *reinterpret_cast<OTM_GLOBAL u32x4*>(out + pl * nn + base) = pz;
      }
    }
  } else {
    // T3[i = n0 + c][l = m0 + r0 .. r0+7] = lo + hi (round 0), T3[n-1-i][l] = lo - hi (round 1)
    const auto out = gp(J.f.mwa);
    for (int round = 0; round < 2; ++round) {
      if (round == 1) __syncthreads();  // round 0's reads of the staged tile are done
      stage(round == 0 ? 1.f : -1.f);
      for (int it = t; it < WT * (WT / 8); it += WTH) {
        const int c = it / (WT / 8), r0 = (it % (WT / 8)) * 8;
        u16 sv[3][8];
#pragma unroll
        for (int q = 0; q < 8; ++q) split3(Ds[(r0 + q) * DL + c], sv[0][q], sv[1][q], sv[2][q]);
        const long b0 = round == 0 ? bofs(n0 + c, m0 + r0, n) : bofs(n - 1 - n0 - c, m0 + r0, n);
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          u32x4 pv;
          pv.x = sv[pl][0] | (unsigned(sv[pl][1]) << 16); pv.y = sv[pl][2] | (unsigned(sv[pl][3]) << 16);
          pv.z = sv[pl][4] | (unsigned(sv[pl][5]) << 16); pv.w = sv[pl][6] | (unsigned(sv[pl][7]) << 16);
          *reinterpret_cast<OTM_GLOBAL u32x4*>(out + pl * nn + b0) = pv;
        }
      }
    }
  }
}

// config 2: four waves of 64 x 64; config 3: eight waves of 64 x 32 (two per SIMD)
template <int ST, int C>
void launch_mstage_w(const JobSet& js, bool gate, int fin, hipStream_t s) {
  constexpr int NW = C == 2 ? 4 : 8;
  const int kind = mixed_kind(ST, C);
  if (js.total[kind] == 0) return;
  hipLaunchKernelGGL((k_mgemm_w<ST, NW>), dim3(unsigned(js.total[kind])), dim3(NW * 64), 0, s, (const Job*)js.jobs,
                     (const int*)js.starts_of(kind), js.count, gate, fin);
  OTM_HIP(hipGetLastError());
}

// config c of the two-group form: tile width MIXED_TN[c]
template <int ST, int C>
void launch_mstage_tn(const JobSet& js, bool gate, int fin, hipStream_t s) {
  constexpr int TN = MIXED_TN[C];
  const int kind = mixed_kind(ST, C);
  if (js.total[kind] == 0) return;
  hipLaunchKernelGGL((k_mgemm<ST, TN>), dim3(unsigned(js.total[kind])), dim3(MNT), 0, s, (const Job*)js.jobs,
                     (const int*)js.starts_of(kind), js.count, gate, fin);
  OTM_HIP(hipGetLastError());
}

// one launch per tile configuration in use (configurations with no job return at once)
template <int ST>
void launch_mstage(const JobSet& js, bool gate, int fin, hipStream_t s) {
  launch_mstage_tn<ST, 0>(js, gate, fin, s);
  launch_mstage_tn<ST, 1>(js, gate, fin, s);
  launch_mstage_w<ST, 2>(js, gate, fin, s);
  launch_mstage_w<ST, 3>(js, gate, fin, s);
}

void launch_mresid(const JobSet& js, bool gate, hipStream_t s) {
  if (js.total[KIND_MRES] == 0) return;
  hipLaunchKernelGGL(k_mresid, dim3(unsigned(js.total[KIND_MRES])), dim3(NT), 0, s, (const Job*)js.jobs,
                     (const int*)js.starts_of(KIND_MRES), js.count, gate);
  OTM_HIP(hipGetLastError());
}

}  // namespace

void launch_mprep(const JobSet& js, bool gate, hipStream_t s) {
  if (js.total[KIND_FLATMX] == 0) return;
  hipLaunchKernelGGL(k_mprep, dim3(unsigned(js.total[KIND_FLATMX])), dim3(NT), 0, s, (const Job*)js.jobs,
                     (const int*)js.starts_of(KIND_FLATMX), js.count, gate);
  OTM_HIP(hipGetLastError());
}

namespace {

void mixed_transform(const JobSet& js, bool gate, int fin, hipStream_t s) {
  launch_mstage<M1>(js, gate, 0, s);
  launch_mstage<M2>(js, gate, 0, s);
  launch_mstage<M3>(js, gate, 0, s);
  launch_mstage<M4>(js, gate, fin, s);
}

}  // namespace

int mixed_tiles(int st, int n, int cfg) {
  const int h = n / 2, tn = MIXED_TN[cfg], tm = MIXED_TM[cfg];
  switch (st) {
    case M1: return (h / (tm / 2)) * (n / tn);
    case M2: return (n / tm) * (n / tn);
    default: return (n / tm) * (h / tn);
  }
}

// Every stage needs h % TN == 0: M3/M4 tile n/2 columns, and an M1/M2 tile
// must not straddle column h, where the A operand switches from the U to the
// W planes.
bool mixed_cfg_valid(int st, int n, int cfg) {
  (void)st;
  return mixed_supported(n) && (n / 2) % MIXED_TN[cfg] == 0;
}

bool mixed_supported(int n) { return n % 64 == 0 && n >= 64; }

void mixed_plan_create(MixedPlan& mp, const DctPlan& p) {
  const int n = p.n, h = n / 2;
  mp.n = n;
  if (!mixed_supported(n) || !p.fold) return;
  const size_t nh = size_t(n) * h;
  std::vector<double> P(nh);
  OTM_HIP(hipMemcpy(P.data(), p.P, nh * sizeof(double), hipMemcpyDeviceToHost));
  // 3-term BF16 split of the FP64 factors (round to nearest even, as __bf16 does)
  auto bf = [](float x) {
    unsigned u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u16(u >> 16);
  };
  auto fl = [](u16 b) {
    const unsigned u = unsigned(b) << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
  };
  std::vector<u16> Pp(3 * nh), Ptp(3 * nh);
  for (int r = 0; r < n; ++r)
    for (int i = 0; i < h; ++i) {
      const double x = P[size_t(r) * h + i];
      const u16 h0 = bf(float(x));
      const double r1 = x - double(fl(h0));
      const u16 h1 = bf(float(r1));
      const u16 h2 = bf(float(r1 - double(fl(h1))));
      const u16 hs[3] = {h0, h1, h2};
      for (int pl = 0; pl < 3; ++pl) {
        Pp[pl * nh + bofs(r, i, h)] = hs[pl];
        Ptp[pl * nh + bofs(i, r, n)] = hs[pl];
      }
    }
  std::vector<double> ck(n);
  for (int r = 0; r < n; ++r) ck[r] = std::cos(M_PI * double(r < h ? 2 * r : 2 * (r - h) + 1) / double(n));
  OTM_HIP(hipMalloc(&mp.ck, n * sizeof(double)));
  OTM_HIP(hipMemcpy(mp.ck, ck.data(), n * sizeof(double), hipMemcpyHostToDevice));
  OTM_HIP(hipMalloc(&mp.Pp, 3 * nh * sizeof(u16)));
  OTM_HIP(hipMalloc(&mp.Ptp, 3 * nh * sizeof(u16)));
  OTM_HIP(hipMemcpy(mp.Pp, Pp.data(), 3 * nh * sizeof(u16), hipMemcpyHostToDevice));
  OTM_HIP(hipMemcpy(mp.Ptp, Ptp.data(), 3 * nh * sizeof(u16), hipMemcpyHostToDevice));
}

void mixed_plan_destroy(MixedPlan& mp) {
  if (mp.Pp) (void)hipFree(mp.Pp);
  if (mp.Ptp) (void)hipFree(mp.Ptp);
  if (mp.ck) (void)hipFree(mp.ck);
  mp = MixedPlan();
}

// one ungated stage (M2, the plain n x n output stage) for kernel timing
void launch_mixed_gemm_stage(const JobSet& js, hipStream_t s) { launch_mstage<M2>(js, false, 0, s); }

namespace {
using MStageFn = void (*)(const JobSet&, bool, int, hipStream_t);
template <int ST>
constexpr std::array<MStageFn, NMIXED> mstage_row() {
  return {launch_mstage_tn<ST, 0>, launch_mstage_tn<ST, 1>, launch_mstage_w<ST, 2>, launch_mstage_w<ST, 3>};
}
const std::array<MStageFn, NMIXED> MSTAGE_TABLE[4] = {mstage_row<M1>(), mstage_row<M2>(), mstage_row<M3>(),
                                                       mstage_row<M4>()};
}  // namespace

// Times each valid tile configuration of every mixed stage over a table of
// the jobs of each grid size on the mixed transform (ungated; M4 writes only
// the x0 scratch) and assigns the fastest to those jobs.  Results are cached per device, grid
// size and number of jobs of that size.
void mixed_tune(JobSet& js, int device, hipStream_t s) {
  static std::mutex mu;
  static std::map<std::array<int, 3>, std::array<int, 4>> cache;
  std::map<int, std::vector<int>> by_size;  // jobs on the mixed transform, by grid size
  for (int p = 0; p < js.count; ++p)
    if (js.jroute[p] == 1) {
      if (!js.host[p].Pp || !mixed_supported(js.host[p].f.n)) return;
      by_size[js.host[p].f.n].push_back(p);
    }
  if (by_size.empty()) return;
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
      for (int st = 0; st < 4; ++st) {
        const int fin = st == M4 ? FIN_X0 : 0;
        best[st] = std::max(0, tune_pick(NMIXED, [&](int c) { return mixed_cfg_valid(st, n, c); },
                                         [&](int c) { MSTAGE_TABLE[st][c](t, false, fin, s); }, s));
      }
      OTM_HIP(hipStreamSynchronize(s));
      jobset_destroy(t);
      std::lock_guard<std::mutex> lock(mu);
      cache[key] = best;
    }
    for (int p : kv.second) js.jmcfg[p] = best;
  }
  jobset_assign(js);
}

// d_hat = M^+ (r - S e0) of every job: mixed transform, then `refine` FP64
// refinement steps x += transform(r' - M0 x); the last stage leaves the
// |d_hat|^2 partials in SITE_DNORM (launch kind KIND_DNORM of these jobs).
// prep = false: the folded planes of r' are already in place (the update
// pass of the previous iteration wrote them).
void launch_mixed_solve(const JobSet& js, int refine, bool gate, hipStream_t s, bool prep) {
  if (prep) launch_mprep(js, gate, s);
  if (refine <= 0) {
    mixed_transform(js, gate, FIN_OUT, s);
    return;
  }
  mixed_transform(js, gate, FIN_X0, s);
  for (int k = 0; k < refine; ++k) {
    launch_mresid(js, gate, s);
    mixed_transform(js, gate, k == refine - 1 ? FIN_ADD_OUT : FIN_ADD_X0, s);
  }
}

}  // namespace otm
