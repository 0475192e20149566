// Stencil, reduction and device-side scalar kernels of the grid transport solver.
//
// Every stencil kernel works on a TI x TJ tile of faces (TJ along the
// contiguous index j) with a one-face halo staged in LDS, derives the
// (TI+1) x (TJ+1) vertex gradients of the tile in LDS, and then evaluates its
// per-face quantity.  Reductions write one partial per workgroup; a
// single-workgroup scalar kernel launched next folds them in a fixed order and
// takes the iteration's scalar decisions (beta, the quartic line search,
// convergence) on the device, so results are deterministic and the host never
// waits inside an iteration.
//
// Arithmetic follows the reference's operation order (the library is built
// with -ffp-contract=off):
//   vertex gradients       otlib/otsolver_2dgrid.cpp:362-423
//   quad area              otlib/otsolver_2dgrid.cpp:425-447
//   residual               otlib/otsolver_2dgrid.cpp:449-465
//   conjugate-Jacobian     otlib/otsolver_2dgrid.cpp:248-261
//   1-D problem parameters otlib/otsolver_2dgrid.cpp:467-539
//   quartic line search    otlib/otsolver_2dgrid.cpp:541-606
//   density adjustment     otlib/otsolver_2dgrid.cpp:47-73

#include <atomic>
#include <cfloat>

#include "otm_internal.h"
#include "reduce.h"
#include "split.h"

namespace otm {
namespace {

constexpr int LDF = TJ + 2;         // face-halo row length
constexpr int NF = (TI + 2) * LDF;  // face-halo values per tile
constexpr int LDV = TJ + 1;         // tile vertex row length
constexpr int NV = (TI + 1) * LDV;  // tile vertices

// Face values fn(face) on [i0-1, i0+TI] x [j0-1, j0+TJ] (0 outside the grid),
// in two phases so that a kernel can issue all its global loads before the
// first LDS store: halo_load fills HR registers per thread, halo_store puts
// them in the LDS tile.
constexpr int HR = (NF + NT - 1) / NT;  // halo values per thread
constexpr int VR = (NV + NT - 1) / NT;  // tile vertices per thread
constexpr int FR = TI * TJ / NT;        // faces per thread
static_assert(TI * TJ % NT == 0, "whole faces per thread");

template <class Fn>
__device__ inline void halo_load(double (&v)[HR], int n, int i0, int j0, Fn fn) {
#pragma unroll
  for (int u = 0; u < HR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int a = t / LDF, b = t - a * LDF;
    const int fi = i0 - 1 + a, fj = j0 - 1 + b;
    v[u] = (t < NF && fi >= 0 && fi < n && fj >= 0 && fj < n) ? fn(long(fi) * n + fj) : 0.0;
  }
}

__device__ inline void halo_store(double* s, const double (&v)[HR]) {
#pragma unroll
  for (int u = 0; u < HR; ++u) {
    const int t = threadIdx.x + u * NT;
    if (t < NF) s[t] = v[u];
  }
}

template <class Fn>
__device__ inline void load_halo(double* s, int n, int i0, int j0, Fn fn) {
  double v[HR];
  halo_load(v, n, i0, j0, fn);
  halo_store(s, v);
}

// Gradient at vertex (vi, vj) = tile-local (a, b) from the face halo s.
// Interior vertices: 2-wide SSE packet order for vj < s2, scalar-tail order
// for vj >= s2, s2 = 2*floor((n-1)/2) (otlib/otsolver_2dgrid.cpp:374-402).
__device__ inline void vertex_grad(const double* s, int n, int a, int b, int vi, int vj, double& gx, double& gy) {
  const double w = double(n);
  const double hw = 0.5 * w;
  const double p00 = s[a * LDF + b];            // face (vi-1, vj-1)
  const double p01 = s[a * LDF + b + 1];        // face (vi-1, vj)
  const double p10 = s[(a + 1) * LDF + b];      // face (vi,   vj-1)
  const double p11 = s[(a + 1) * LDF + b + 1];  // face (vi,   vj)
  const bool ei = (vi == 0) || (vi == n);
  const bool ej = (vj == 0) || (vj == n);
  if (ei && ej) {
    gx = 0.; gy = 0.;
  } else if (vj == 0) {
    gx = w * (p11 - p01); gy = 0.;
  } else if (vj == n) {
    gx = w * (p10 - p00); gy = 0.;
  } else if (vi == 0) {
    gx = 0.; gy = w * (p11 - p10);
  } else if (vi == n) {
    gx = 0.; gy = w * (p01 - p00);
  } else if (vj < ((n - 1) / 2) * 2) {
    gx = hw * ((p10 + p11) - (p00 + p01));
    gy = hw * ((p01 + p11) - (p00 + p10));
  } else {
    gx = hw * (((p10 + p11) - p00) - p01);
    gy = hw * (((p01 + p11) - p00) - p10);
  }
}

__device__ inline void tile_grads(const double* s, double* gx, double* gy, int n, int i0, int j0) {
  for (int t = threadIdx.x; t < NV; t += NT) {
    const int a = t / LDV, b = t - a * LDV;
    const int vi = i0 + a, vj = j0 + b;
    double x = 0., y = 0.;
    if (vi <= n && vj <= n) vertex_grad(s, n, a, b, vi, vj, x, y);
    gx[t] = x;
    gy[t] = y;
  }
}

// Vertex gradients of two face fields whose halos sa, sb sit in LDS, into
// the four arrays gx, gy (of sa) and hx, hy (of sb) over the same LDS: the
// gradients go through registers, so the halos may lie under the arrays.
// Starts after a barrier that made both halos visible; ends with one.
__device__ inline void two_tile_grads(const double* sa, const double* sb, double* gx, double* gy, double* hx,
                                      double* hy, int n, int i0, int j0) {
  double gv[VR][4];
#pragma unroll
  for (int u = 0; u < VR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int a = t / LDV, b = t - a * LDV;
    const int vi = i0 + a, vj = j0 + b;
    gv[u][0] = gv[u][1] = gv[u][2] = gv[u][3] = 0.;
    if (t < NV && vi <= n && vj <= n) {
      vertex_grad(sa, n, a, b, vi, vj, gv[u][0], gv[u][1]);
      vertex_grad(sb, n, a, b, vi, vj, gv[u][2], gv[u][3]);
    }
  }
  __syncthreads();  // every halo read is done: the arrays take their place
#pragma unroll
  for (int u = 0; u < VR; ++u) {
    const int t = threadIdx.x + u * NT;
    if (t < NV) {
      gx[t] = gv[u][0];
      gy[t] = gv[u][1];
      hx[t] = gv[u][2];
      hy[t] = gv[u][3];
    }
  }
  __syncthreads();
}

// half the cross product of the quad diagonals, tile-local face (a, b)
__device__ inline double quad_area(const double* gx, const double* gy, int a, int b, double e) {
  const int v00 = a * LDV + b, v10 = (a + 1) * LDV + b, v01 = a * LDV + b + 1, v11 = (a + 1) * LDV + b + 1;
  return 0.5 * ((gx[v11] - gx[v00] + e) * (gy[v01] - gy[v10] + e) -
                (gy[v11] - gy[v00] + e) * (gx[v01] - gx[v10] - e));
}

// Quartic coefficients of tile-local face (a, b): with cached gradients
// G0 = (hx, hy) and direction gradients (gx, gy), r(psi + t d) = r + b t + a t^2
// (otsolver_2dgrid.cpp:496-537).  Shared by the line-search pass and the
// update pass so that both evaluate exactly the same expressions.
__device__ inline void quartic_ab(const double* gx, const double* gy, const double* hx, const double* hy, int a,
                                  int b, double e, double& av, double& bv) {
  const int v00 = a * LDV + b, v10 = (a + 1) * LDV + b, v01 = a * LDV + b + 1, v11 = (a + 1) * LDV + b + 1;
  const double d0ax = hx[v11] - hx[v00] + e;
  const double d0ay = hy[v11] - hy[v00] + e;
  const double d0bx = hx[v01] - hx[v10] - e;
  const double d0by = hy[v01] - hy[v10] + e;
  const double ddax = gx[v11] - gx[v00];
  const double dday = gy[v11] - gy[v00];
  const double ddbx = gx[v01] - gx[v10];
  const double ddby = gy[v01] - gy[v10];
  av = 0.5 * (ddax * ddby - dday * ddbx);
  bv = 0.5 * (ddax * d0by - dday * d0bx + d0ax * ddby - d0ay * ddbx);
}

// ---------------------------------------------------------------------------
// Exact quartic line search on the device.
// e(t) = z0 + z1 t + z2 t^2 + z3 t^3 + z4 t^4; the stationary points are the
// real roots of e'(t)/w3 = t^3 + c2 t^2 + c1 t + c0 (w = (z1, 2z2, 3z3, 4z4),
// c = w/w3 as the companion matrix of the reference).  Roots come from the
// closed form of the depressed cubic (Cardano for one real root, Viete's
// trigonometric form for three) and are polished by Newton steps.  Among the
// real roots in [-10, 10] the one of least e(t) wins (first on ties); none
// found -> alpha = 0.5.
// ---------------------------------------------------------------------------
__device__ inline double cubic_eval(double t, double c2, double c1, double c0) { return ((t + c2) * t + c1) * t + c0; }

__device__ inline double newton_polish(double t, double c2, double c1, double c0) {
  for (int k = 0; k < 4; ++k) {
    const double f = cubic_eval(t, c2, c1, c0);
    const double df = (3. * t + 2. * c2) * t + c1;
    if (f == 0. || df == 0.) break;
    const double tn = t - f / df;
    if (!(fabs(cubic_eval(tn, c2, c1, c0)) < fabs(f))) break;
    t = tn;
  }
  return t;
}

__device__ inline int cubic_roots_dev(double c2, double c1, double c0, double* roots) {
  const double sh = -c2 / 3.;
  const double p = c1 - c2 * c2 / 3.;
  const double q = (2. * c2 * c2 * c2) / 27. - (c2 * c1) / 3. + c0;
  const double hq = 0.5 * q, tp = p / 3.;
  const double disc = hq * hq + tp * tp * tp;
  int k = 0;
  if (disc > 0.) {
    const double sq = sqrt(disc);
    const double A = -copysign(cbrt(fabs(hq) + sq), hq);
    const double B = (A != 0.) ? -tp / A : 0.;
    roots[k++] = A + B + sh;
  } else if (p == 0.) {
    roots[k++] = sh;  // triple root
  } else {
    const double m = 2. * sqrt(-tp);
    double arg = (3. * q) / (p * m);
    arg = fmin(1., fmax(-1., arg));
    const double phi = acos(arg) / 3.;
    const double two_pi_3 = 2.0943951023931954923;  // 2*pi/3
    roots[k++] = m * cos(phi) + sh;
    roots[k++] = m * cos(phi - two_pi_3) + sh;
    roots[k++] = m * cos(phi - 2. * two_pi_3) + sh;
  }
  for (int i = 0; i < k; ++i) roots[i] = newton_polish(roots[i], c2, c1, c0);
  return k;
}

__device__ inline double pick_step(const double* z, double& emin) {
  const double w0 = z[1], w1 = 2. * z[2], w2 = 3. * z[3], w3 = 4. * z[4];
  const double c0 = w0 / w3, c1 = w1 / w3, c2 = w2 / w3;
  double roots[3];
  int nr = 0;
  if (isfinite(c0) && isfinite(c1) && isfinite(c2)) nr = cubic_roots_dev(c2, c1, c0, roots);
  bool found = false;
  double alpha = 0., rmin = 0.;
  for (int k = 0; k < nr; ++k) {
    const double t = roots[k];
    if (t >= -10. && t <= 10.) {
      const double r = z[0] + t * (z[1] + t * (z[2] + t * (z[3] + t * z[4])));
      if (!found || r < rmin) { found = true; alpha = t; rmin = r; }
    }
  }
  if (!found) {
    alpha = 0.5;
    rmin = z[0] + alpha * (z[1] + alpha * (z[2] + alpha * (z[3] + alpha * z[4])));
  }
  emin = rmin;
  return alpha;
}

// ---------------------------------------------------------------------------
// kernels of the solve, over a job table (see Job / JobSet)
// ---------------------------------------------------------------------------
#define JOB_ARGS const Job* __restrict__ jobs, const int* __restrict__ starts, int count

struct Loc {
  int p;   // job
  int t;   // workgroup within the job
  int nb;  // workgroups of the job in this launch
};

__device__ inline Loc locate(const int* __restrict__ starts, int count) {
  const int b = blockIdx.x;
  if (count == 1) return {0, b, int(gridDim.x)};  // one job: no table lookup
  const int p = job_of(starts, count, b);
  return {p, b - starts[p], starts[p + 1] - starts[p]};
}

// Stencil launches run one workgroup per tile; workgroup w of a job takes
// tile band_tile(w, T) of its T tiles.  Workgroups are dealt to the 8 XCDs
// round-robin, so XCD x gets tiles [x T/8, (x + 1) T/8): a band of
// consecutive tile rows, in which the left and right neighbours of a tile
// (whose edge faces it reads as its halo) share its L2, as do the upper and
// lower ones except at the band's edges.  (With the row-major order an
// XCD got every eighth tile column, and no horizontal neighbour.)  A
// permutation of [0, T), the identity when T % 8 != 0; partials are written
// at the tile index, so the fold order does not depend on it.
__device__ inline int band_tile(int w, int T) { return (T & 7) ? w : (w & 7) * (T >> 3) + (w >> 3); }
__device__ inline Loc locate_tile(const int* __restrict__ starts, int count) {
  Loc L = locate(starts, count);
  L.t = band_tile(L.t, L.nb);
  return L;
}

// tile origin of local stencil tile t (row-major over the tiles_j(n) x tiles_i(n) grid)
__device__ inline void tile_origin(int t, int n, int& i0, int& j0) {
  const int tj = (n + TJ - 1) / TJ;
  i0 = (t / tj) * TI;
  j0 = (t % tj) * TJ;
}

__global__ __launch_bounds__(NT) void k_fill(double* p, long count, double v) {
  for (long i = long(blockIdx.x) * NT + threadIdx.x; i < count; i += long(gridDim.x) * NT) p[i] = v;
}

// psi = 0, both search directions = 0 (otsolver_2dgrid.cpp:126-140); the
// cached vertex gradients of psi (zero at psi = 0) are not stored: the
// passes that use them form grad psi from psi
__global__ __launch_bounds__(NT) void k_init_fields(JOB_ARGS) {
  const Loc L = locate(starts, count);
  const Fields& f = jobs[L.p].f;
  const long stride = long(L.nb) * NT;
  const auto psi = gp(f.psi), d0 = gp(f.d[0]), d1 = gp(f.d[1]);
  for (long i = long(L.t) * NT + threadIdx.x; i < f.F; i += stride) {
    psi[i] = 0.;
    d0[i] = 0.;
    d1[i] = 0.;
  }
}

// density adjustment (otsolver_2dgrid.cpp:47-73), pass 1: sum / min / max
__global__ __launch_bounds__(NT) void k_adjust_reduce(JOB_ARGS) {
  const Loc L = locate(starts, count);
  const Job& J = jobs[L.p];
  const auto in = gp(J.dens_in);
  const long F = J.f.F;
  double s[1] = {0.};
  double mn = DBL_MAX, mx = -DBL_MAX;
  for (long i = long(L.t) * NT + threadIdx.x; i < F; i += long(L.nb) * NT) {
    const double v = in[i];
    s[0] += v;
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mn = fmin(mn, __shfl_down(mn, off));
    mx = fmax(mx, __shfl_down(mx, off));
  }
  __shared__ double mm[2][NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { mm[0][w] = mn; mm[1][w] = mx; }
  block_sum<1>(s);  // synchronises after the mm writes
  if (threadIdx.x == 0) {
    for (int k = 1; k < NT / 64; ++k) { mm[0][0] = fmin(mm[0][0], mm[0][k]); mm[1][0] = fmax(mm[1][0], mm[1][k]); }
    const auto partials = gp(J.f.site(SITE_ADJUST));
    partials[L.t] = s[0];
    partials[L.nb + L.t] = mm[0][0];
    partials[2 * L.nb + L.t] = mm[1][0];
  }
}

// pass 2 (one workgroup per job): integral, ratio cap -> scale and shift
__global__ __launch_bounds__(NT) void k_adjust_scalar(JOB_ARGS, double max_ratio) {
  const int p = blockIdx.x;
  const Fields& f = jobs[p].f;
  const int nb = starts[p + 1] - starts[p];
  const auto partials = gp(f.site(SITE_ADJUST));
  const double A = 1.0 / (double(f.n) * double(f.n));
  double tot[1] = {0.};
  double gmn = DBL_MAX, gmx = -DBL_MAX;
  for (int j = threadIdx.x; j < nb; j += NT) {
    tot[0] += partials[j];
    gmn = fmin(gmn, partials[nb + j]);
    gmx = fmax(gmx, partials[2 * nb + j]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    gmn = fmin(gmn, __shfl_down(gmn, off));
    gmx = fmax(gmx, __shfl_down(gmx, off));
  }
  __shared__ double mm[2][NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { mm[0][w] = gmn; mm[1][w] = gmx; }
  block_sum<1>(tot);
  if (threadIdx.x == 0) {
    for (int k = 1; k < NT / 64; ++k) { mm[0][0] = fmin(mm[0][0], mm[0][k]); mm[1][0] = fmax(mm[1][0], mm[1][k]); }
    const double I = tot[0] * A;
    const double nmn = mm[0][0] / I, nmx = mm[1][0] / I;  // division by I > 0 is monotone
    const double ratio = fmin(max_ratio, nmx / nmn);
    const double cand = (ratio * nmn - nmx) / (1. - ratio);
    const auto st = gp(f.state);
    st->dens_scale = I;
    st->dens_shift = (0. < cand) ? cand : 0.;  // std::max(0., cand), NaN -> 0
    st->max_ratio = max_ratio;
  }
}

__global__ __launch_bounds__(NT) void k_adjust_apply(JOB_ARGS) {
  const Loc L = locate(starts, count);
  const Job& J = jobs[L.p];
  const auto in = gp(J.dens_in);
  const auto rho = gp(J.f.rho);
  const auto st = gp((const SolveState*)J.f.state);
  const double I = st->dens_scale, b = st->dens_shift;
  for (long i = long(L.t) * NT + threadIdx.x; i < J.f.F; i += long(L.nb) * NT) rho[i] = (in[i] / I + b) / (1. + b);
}

// initial residual at psi = 0: r_0 and the partials of |r|^2, sum r
__global__ __launch_bounds__(NT) void k_initial_residual(JOB_ARGS) {
  const Loc L = locate_tile(starts, count);
  const Fields& f = jobs[L.p].f;
  const int n = f.n;
  __shared__ double gx[NV], gy[NV];
  __shared__ double s[NF];
  int i0, j0;
  tile_origin(L.t, n, i0, j0);
  load_halo(s, n, i0, j0, [](long) { return 0.0; });
  __syncthreads();
  tile_grads(s, gx, gy, n, i0, j0);
  __syncthreads();
  const double e = 1. / double(n), A = 1.0 / (double(n) * double(n));
  double acc[2] = {0., 0.};
  for (int t = threadIdx.x; t < TI * TJ; t += NT) {
    const int a = t / TJ, b = t % TJ;
    const int fi = i0 + a, fj = j0 + b;
    if (fi < n && fj < n) {
      const long fc = long(fi) * n + fj;
      const double rv = quad_area(gx, gy, a, b, e) - A * gp(f.rho)[fc];
      gp(f.r)[fc] = rv;
      acc[0] += rv * rv;
      acc[1] += rv;
    }
  }
  write_partials_at<2>(acc, f.site(SITE_RESID), L.t, L.nb);
}

// A job still far from its stop: residual above NEAR_FACTOR x threshold and
// more than one long replay of iterations below max_iter (the host queues
// long replays while some job is far, single iterations after that).
__device__ inline bool far_from_stop(double e, double threshold, int it, int max_iter) {
  return e > NEAR_FACTOR * threshold && it + ITERS_PER_GRAPH < max_iter;
}

// solver state initialisation (otsolver_2dgrid.cpp:126-160), one workgroup per job.
// flags (one-job tables, else null): the host's {open, far} pair in mapped
// pinned memory; the scalar kernels write it where either changes, so the
// host reads the job's progress without a count kernel per replay.
__global__ __launch_bounds__(NT) void k_scalar_init(JOB_ARGS, int max_iter, double threshold, int cj, int* flags) {
  const int p = blockIdx.x;
  const Fields& f = jobs[p].f;
  double acc[2];
  fold_partials<2>(acc, f.site(SITE_RESID), starts[p + 1] - starts[p]);
  if (threadIdx.x == 0) {
    const auto st = gp(f.state);
    const double A = 1.0 / (double(f.n) * double(f.n));
    const double e0 = acc[0] / A;
    st->e = e0;
    st->e0 = e0;
    st->sum_r = acc[1];
    st->alpha = 0.;
    st->beta = 0.;
    st->dnorm = 0.;
    st->it = 0;
    st->max_iter = max_iter;
    st->threshold = threshold;
    st->cj = cj;
    st->update = 0;
    st->use_prev = 0;
    const bool run = (0 < max_iter) && (e0 > threshold);
    st->done = run ? 0 : 1;
    st->exit_reason = run ? 0 : 1;
    if (flags) {
      flags[0] = run ? 1 : 0;
      flags[1] = run && far_from_stop(e0, threshold, 0, max_iter) ? 1 : 0;
    }
  }
}

// conjugate-Jacobian beta, pass 1: the partials of Jd.(r_eps - r_k) and
// Jd.Jd (otsolver_2dgrid.cpp:248-261), r_eps the residual at psi - eps d_hat
// (eps = alpha/2).  The area is quadratic in the vertex gradients, which are
// linear in the potential, so with the quartic coefficients (a, b) of d_hat
// at psi (the line search's, otsolver_2dgrid.cpp:467-539)
// r(psi + t d_hat) - r(psi) = b t + a t^2 exactly, and r_eps - r_k is that at
// t = -eps: the pass reads psi's and d_hat's halos and Jd, not rho and r.
// (The reference evaluates r_eps afresh and subtracts the running r_k; the
// two differ by the running residual's rounding drift.)
__global__ __launch_bounds__(NT) void k_beta(JOB_ARGS) {
  const Loc L = locate_tile(starts, count);
  const Fields& f = jobs[L.p].f;
  const auto st = gp((const SolveState*)f.state);
  // one job: the tile's reads are issued before the gate is read (the two
  // latencies overlap); in a batch the gate comes first
  const bool early = count == 1;
  if (!early && (st->done || st->it < 1 || !st->cj)) return;
  const int n = f.n;
  const auto psi = gp(f.psi_buf(st->it));
  const auto dhat = gp(f.dhat);
  const auto jd = gp(f.jd);
  // LDS as in the line-search pass: the halos of psi (sp) and d_hat (s)
  // under the gradients gx, gy (d_hat) and hx, hy (psi)
  __shared__ double lds[4 * NV];
  double* const gx = lds;
  double* const gy = lds + NV;
  double* const hx = lds + 2 * NV;
  double* const hy = lds + 3 * NV;
  double* const sp = gx;
  double* const s = hx;
  const double eps = st->alpha / 2.;
  int i0, j0;
  tile_origin(L.t, n, i0, j0);
  // all global loads of the tile first: the two halos, then Jd of its faces
  double pa[HR], pb[HR];
#pragma unroll
  for (int u = 0; u < HR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int a = t / LDF, b = t - a * LDF;
    const int fi = i0 - 1 + a, fj = j0 - 1 + b;
    const bool in = t < NF && fi >= 0 && fi < n && fj >= 0 && fj < n;
    const long fc = long(fi) * n + fj;
    pa[u] = in ? psi[fc] : 0.;
    pb[u] = in ? dhat[fc] : 0.;
  }
  double jdv[FR];
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int fi = i0 + t / TJ, fj = j0 + t % TJ;
    jdv[u] = (fi < n && fj < n) ? jd[long(fi) * n + fj] : 0.;
  }
  if (early && (st->done || st->it < 1 || !st->cj)) return;
  halo_store(s, pb);
  halo_store(sp, pa);
  __syncthreads();
  two_tile_grads(s, sp, gx, gy, hx, hy, n, i0, j0);
  const double e = 1. / double(n);
  const double t = -eps;
  double acc[2] = {0., 0.};
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int q = threadIdx.x + u * NT;
    const int a = q / TJ, b = q % TJ;
    const int fi = i0 + a, fj = j0 + b;
    if (fi < n && fj < n) {
      double av, bv;
      quartic_ab(gx, gy, hx, hy, a, b, e, av, bv);
      const double j = jdv[u];
      acc[0] += j * (av * (t * t) + bv * t);  // Jd (r_eps - r_k)
      acc[1] += j * j;
    }
  }
  write_partials_at<2>(acc, f.site(SITE_BETA), L.t, L.nb);
}

// pass 2 (one workgroup per job): early stop on |d_hat| (otsolver_2dgrid.cpp:182-183),
// then beta = max(-1, Jd.(r_eps - r_k) / |Jd|^2 / eps * alpha) and the choice
// d = d_hat + beta d_prev (it >= 1, conjugate-Jacobian) or d = d_hat.
// dstarts: workgroup prefix of the launch that wrote the |d_hat|^2 partials.
// Wide form (jobs of more than NT * FOLD_BATCH stencil tiles): the
// partials are folded by BETA_GROUPS groups of NT threads, with the same
// result as NT threads where there are at most that many.
constexpr int BETA_GROUPS = 4;
template <bool WIDE>
__global__ __launch_bounds__(WIDE ? BETA_GROUPS * NT : NT) void k_scalar_beta(JOB_ARGS, const int* __restrict__ dstarts, int* flags) {
  const int p = blockIdx.x;
  const Fields& f = jobs[p].f;
  const auto st = gp(f.state);
  // the state is read together with the partials (one memory round trip);
  // partials that no pass wrote this iteration are folded and discarded
  const int done = st->done, it = st->it, cj = st->cj;
  const double alpha = st->alpha;
  double v[3];  // |d_hat|^2, Jd.(r_eps - r_k), Jd.Jd: all loads issued before the sum
  {
    double dn[1] = {0.}, bd[2] = {0., 0.};
    if (WIDE) {
      fold_share_grouped<1, BETA_GROUPS>(dn, f.site(SITE_DNORM), dstarts[p + 1] - dstarts[p]);
      fold_share_grouped<2, BETA_GROUPS>(bd, f.site(SITE_BETA), starts[p + 1] - starts[p]);
    } else {
      fold_share<1>(dn, f.site(SITE_DNORM), dstarts[p + 1] - dstarts[p]);
      fold_share<2>(bd, f.site(SITE_BETA), starts[p + 1] - starts[p]);
    }
    v[0] = dn[0];
    v[1] = bd[0];
    v[2] = bd[1];
  }
  if (WIDE)
    group_sum<3, BETA_GROUPS>(v);
  else
    block_sum<3>(v);
  if (threadIdx.x != 0) return;
  st->update = 0;  // the previous iteration's update has run
  if (done) return;
  const double dnorm = sqrt(v[0]);
  st->dnorm = dnorm;
  if (dnorm <= 2 * DBL_MIN) {
    st->done = 1;
    st->exit_reason = 2;
    if (flags) flags[0] = 0;
    return;
  }
  const bool conj = it >= 1 && cj;
  if (!conj) {
    st->use_prev = 0;  // beta keeps its value (printed as is)
    return;
  }
  const double eps = alpha / 2.;
  const double cand = v[1] / v[2] / eps * alpha;
  st->beta = (-1. < cand) ? cand : -1.;  // std::max(-1., cand)
  st->use_prev = 1;
}

// direction d = d_hat (+ beta d_prev), vertex gradients of d, quartic
// coefficients a, b per quad (r(psi + t d) = r + b t + a t^2), and the
// partials of b.r, a.r, b.b, a.b, a.a, sum a, sum b   (otsolver_2dgrid.cpp:467-561).
// a and b are not stored: the update pass evaluates them again from the same
// inputs (d and psi before its update), which costs less than their HBM round trip.
__global__ __launch_bounds__(NT) void k_dir_coef(JOB_ARGS) {
  const Loc L = locate_tile(starts, count);
  const Fields& f = jobs[L.p].f;
  const auto st = gp((const SolveState*)f.state);
  if (st->done) return;
  const int n = f.n;
  // LDS (18.7 KB per workgroup: eight resident per CU, every tile of a
  // 1024^2 grid at once): the halos of psi (sp) and of the direction (s),
  // then, over them, the vertex gradients of the direction (gx, gy) and of
  // psi (hx, hy; the reference's cached G0, otsolver_2dgrid.cpp:473-478,
  // which equals grad psi: it starts at 0 with psi = 0 and takes the same
  // alpha-steps)
  __shared__ double lds[4 * NV];
  static_assert(NF <= 2 * NV, "a halo fits under two gradient arrays");
  double* const gx = lds;
  double* const gy = lds + NV;
  double* const hx = lds + 2 * NV;
  double* const hy = lds + 3 * NV;
  double* const sp = gx;
  double* const s = hx;
  const int k = st->it;
  const auto dhat = gp(f.dhat);
  const auto dprev = gp(f.d[k & 1]);
  const auto dnew = gp(f.d[(k + 1) & 1]);
  const auto psi = gp(f.psi_buf(k));
  const auto r = gp(f.r);
  const bool use_prev = st->use_prev != 0;
  const double beta = st->beta;
  int i0, j0;
  tile_origin(L.t, n, i0, j0);
  // all global loads of the tile first: direction halo (d_hat and d_prev
  // kept apart, combined at the LDS store, so no arithmetic waits on a load
  // before the others are issued), psi halo, residual of its faces
  double da[HR], db[HR], pa[HR];
  bool hin[HR];
#pragma unroll
  for (int u = 0; u < HR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int a = t / LDF, b = t - a * LDF;
    const int fi = i0 - 1 + a, fj = j0 - 1 + b;
    hin[u] = t < NF && fi >= 0 && fi < n && fj >= 0 && fj < n;
    const long fc = long(fi) * n + fj;
    da[u] = hin[u] ? dhat[fc] : 0.;
    db[u] = (hin[u] && use_prev) ? dprev[fc] : 0.;
    pa[u] = hin[u] ? psi[fc] : 0.;
  }
  double rvv[FR];
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int fi = i0 + t / TJ, fj = j0 + t % TJ;
    rvv[u] = (fi < n && fj < n) ? r[long(fi) * n + fj] : 0.;
  }
#pragma unroll
  for (int u = 0; u < HR; ++u) {
    const int t = threadIdx.x + u * NT;
    if (t < NF) {
      s[t] = !hin[u] ? 0. : use_prev ? da[u] + beta * db[u] : da[u];
      sp[t] = pa[u];
    }
  }
  __syncthreads();
  double dv[FR];  // direction at the tile's own faces
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int t = threadIdx.x + u * NT;
    dv[u] = s[(t / TJ + 1) * LDF + t % TJ + 1];
  }
  two_tile_grads(s, sp, gx, gy, hx, hy, n, i0, j0);
  const double e = 1. / double(n);
  double acc[7] = {0., 0., 0., 0., 0., 0., 0.};
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int a = t / TJ, b = t % TJ;
    const int fi = i0 + a, fj = j0 + b;
    if (fi < n && fj < n) {
      const long fc = long(fi) * n + fj;
      dnew[fc] = dv[u];
      double av, bv;
      quartic_ab(gx, gy, hx, hy, a, b, e, av, bv);
      const double rv = rvv[u];
      acc[0] += bv * rv;
      acc[1] += av * rv;
      acc[2] += bv * bv;
      acc[3] += av * bv;
      acc[4] += av * av;
      acc[5] += av;
      acc[6] += bv;
    }
  }
  write_partials_at<7>(acc, f.site(SITE_COEF), L.t, L.nb);
}

// pass 2 (one workgroup per job): z = [e A, 2 b.r, b.b + 2 a.r, 2 a.b, a.a], the
// step alpha, e(alpha)/A, traces, the iteration count and the stop test
// (otsolver_2dgrid.cpp:556-605, :160); sum r_{k+1} = alpha^2 sum a + alpha sum b + sum r_k.
// Wide form: two groups of NT threads fold the partials (the step logic of
// thread 0 needs more registers than four groups leave a thread).
constexpr int STEP_GROUPS = 2;
template <bool WIDE>
__global__ __launch_bounds__(WIDE ? STEP_GROUPS * NT : NT) void k_scalar_step(JOB_ARGS, int* flags) {
  const int p = blockIdx.x;
  const Fields& f = jobs[p].f;
  const auto st = gp(f.state);
  // the state is read together with the partials (one memory round trip);
  // a finished job folds them too and discards the result
  const int done = st->done, k = st->it, max_iter = st->max_iter;
  const double e_prev = st->e, sum_r = st->sum_r, beta = st->beta, threshold = st->threshold;
  double acc[7];
  if (WIDE) {
    for (int q = 0; q < 7; ++q) acc[q] = 0.;
    fold_share_grouped<7, STEP_GROUPS>(acc, f.site(SITE_COEF), starts[p + 1] - starts[p]);
    group_sum<7, STEP_GROUPS>(acc);
  } else {
    fold_partials<7>(acc, f.site(SITE_COEF), starts[p + 1] - starts[p]);
  }
  if (threadIdx.x == 0 && !done) {
    const double A = 1.0 / (double(f.n) * double(f.n));
    double z[5];
    z[0] = e_prev * A;
    z[1] = 2. * acc[0];
    z[2] = acc[2] + 2 * acc[1];
    z[3] = 2. * acc[3];
    z[4] = acc[4];
    double emin;
    const double alpha = pick_step(z, emin);
    const double e = emin / A;
    for (int q = 0; q < 5; ++q) st->z[q] = z[q];
    st->alpha = alpha;
    st->e = e;
    st->sum_r = acc[5] * (alpha * alpha) + acc[6] * alpha + sum_r;
    if (k < f.trace_cap) {
      const auto tr = gp(f.trace);
      tr[k] = e;
      tr[f.trace_cap + k] = alpha;
      tr[2 * f.trace_cap + k] = beta;
    }
    const int it = k + 1;
    st->it = it;
    st->update = 1;
    const bool stop = !(it < max_iter && e > threshold);
    if (stop) {
      st->done = 1;
      st->exit_reason = 1;
    }
    if (flags) {
      if (stop) flags[0] = 0;
      const bool far_prev = far_from_stop(e_prev, threshold, k, max_iter);
      const bool far_now = !stop && far_from_stop(e, threshold, it, max_iter);
      if (far_now != far_prev) flags[1] = far_now ? 1 : 0;
    }
  }
}

// Inputs of the update of one tile, loaded into registers: the direction's
// and psi's halos, r of its faces (the update overwrites psi and r, so a
// workgroup may load the inputs of its next tile before it updates the
// current one).
struct UpdateIn {
  double hv[HR];   // direction halo
  double hp[HR];   // psi halo
  double rkv[FR];  // r of the tile's faces
};

__device__ inline void update_load(const Fields& f, int k, int i0, int j0, UpdateIn& R) {
  const int n = f.n;
  const auto d = gp(f.d[(k + 1) & 1]);
  const auto psi = gp(f.psi_buf(k)), r = gp(f.r);
  halo_load(R.hv, n, i0, j0, [&](long fc) { return d[fc]; });
  halo_load(R.hp, n, i0, j0, [&](long fc) { return psi[fc]; });
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int fi = i0 + t / TJ, fj = j0 + t % TJ;
    const bool in = fi < n && fj < n;
    R.rkv[u] = in ? r[long(fi) * n + fj] : 0.;
  }
}

// line-search update of the tile at (i0, j0) from its loaded inputs: psi +=
// alpha d, r <- a alpha^2 + b alpha + r (otsolver_2dgrid.cpp:599-604) and
// Jd = r_{k+1} - r_k for the next beta; k = completed iterations - 1.  The
// quartic coefficients a, b come from d and grad psi before the update, as
// in the line-search pass (the reference's g0 += alpha grad d is the
// gradient of the updated psi, which the next passes take from psi).
// rn: the new residual of this thread's faces (0 outside).  Starts by
// writing the workgroup's LDS: a second call needs a barrier after the first.
__device__ inline void update_apply(const Fields& f, int k, double alpha, int i0, int j0, const UpdateIn& R,
                                    double (&rn)[FR]) {
  const int n = f.n;
  // LDS as in the line-search pass: the halos of psi (sp) and of the
  // direction (s) under the gradients gx, gy (direction) and hx, hy (psi)
  __shared__ double lds[4 * NV];
  double* const gx = lds;
  double* const gy = lds + NV;
  double* const hx = lds + 2 * NV;
  double* const hy = lds + 3 * NV;
  double* const sp = gx;
  double* const s = hx;
  const auto psin = gp(f.psi_buf(k + 1)), r = gp(f.r), jd = gp(f.jd);
  halo_store(s, R.hv);
  halo_store(sp, R.hp);
  __syncthreads();
  double dv[FR], pv[FR];  // direction and psi at the tile's own faces
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int h = (t / TJ + 1) * LDF + t % TJ + 1;
    dv[u] = s[h];
    pv[u] = sp[h];
  }
  two_tile_grads(s, sp, gx, gy, hx, hy, n, i0, j0);
  const double e = 1. / double(n);
#pragma unroll
  for (int u = 0; u < FR; ++u) {
    const int t = threadIdx.x + u * NT;
    const int fa = t / TJ, fb = t % TJ;
    const int fi = i0 + fa, fj = j0 + fb;
    rn[u] = 0.;
    if (fi < n && fj < n) {
      const long fc = long(fi) * n + fj;
      psin[fc] = pv[u] + alpha * dv[u];
      double av, bv;
      quartic_ab(gx, gy, hx, hy, fa, fb, e, av, bv);
      const double rk = R.rkv[u];
      const double rnv = av * (alpha * alpha) + bv * alpha + rk;
      r[fc] = rnv;
      jd[fc] = rnv - rk;
      rn[u] = rnv;
    }
  }
}

// eight waves per SIMD: every tile of a 1024^2 grid resident at once
__global__ __launch_bounds__(NT, 8) void k_update(JOB_ARGS) {
  const Loc L = locate_tile(starts, count);
  const Fields& f = jobs[L.p].f;
  const auto st = gp((const SolveState*)f.state);
  if (!st->update) return;
  int i0, j0;
  tile_origin(L.t, f.n, i0, j0);
  const int k = st->it - 1;
  UpdateIn R;
  update_load(f, k, i0, j0, R);
  double rn[FR];
  update_apply(f, k, st->alpha, i0, j0, R, rn);
}

// The update pass of the mixed form (n % 64 == 0): one workgroup updates a
// tile and its mirror across column n/2 (tile columns c and T-1-c, T = n/TJ;
// the middle tile column alone when T is odd), then writes the next Laplacian
// solve's right-hand side: the columns of r' = r - (sum r) e0 folded (U =
// r'(:, j) + r'(:, n-1-j), W = difference) and split into the six BF16 planes
// (blocked layout), the same values the separate fold kernel writes from r.
__global__ __launch_bounds__(NT) void k_update_fold(JOB_ARGS) {
  const Loc L = locate(starts, count);
  const Job& J = jobs[L.p];
  const Fields& f = J.f;
  const auto st = gp((const SolveState*)f.state);
  if (!st->update) return;
  const int n = f.n, h = n / 2, T = n / TJ, P = (T + 1) / 2;
  const int i0 = (L.t / P) * TI, cl = L.t % P, cr = T - 1 - cl;
  const int k = st->it - 1;
  const double alpha = st->alpha;
  __shared__ double rl[TI * TJ], rr[TI * TJ];
  // both tiles' inputs are loaded before the first update (the mirror's
  // loads overlap the first tile's work)
  UpdateIn RL, RR;
  update_load(f, k, i0, cl * TJ, RL);
  if (cr != cl) update_load(f, k, i0, cr * TJ, RR);
  double rn[FR];
  update_apply(f, k, alpha, i0, cl * TJ, RL, rn);
#pragma unroll
  for (int u = 0; u < FR; ++u) rl[threadIdx.x + u * NT] = rn[u];
  if (cr != cl) {
    __syncthreads();  // the tile's LDS is reused
    update_apply(f, k, alpha, i0, cr * TJ, RR, rn);
#pragma unroll
    for (int u = 0; u < FR; ++u) rr[threadIdx.x + u * NT] = rn[u];
  }
  __syncthreads();
  // fold: face (a, b) of the left tile, column j = cl*TJ + b < h, pairs with
  // column n-1-j: face (a, TJ-1-b) of the right tile (of the same tile when
  // it is the middle one); four consecutive columns per thread
  const double* mir = cr != cl ? rr : rl;
  const int cols = cr != cl ? TJ : TJ / 2;
  const int t = threadIdx.x, a = t / (TJ / 4), b0 = (t % (TJ / 4)) * 4;
  if (a < TI && b0 < cols) {
    const double S = st->sum_r;
    const long nh = long(n) * h;
    const auto out = gp(J.f.mrhs);
    const int i = i0 + a, j = cl * TJ + b0;
    u16 uu[3][4], ww[3][4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double va = rl[a * TJ + b0 + e];
      if (i == 0 && j + e == 0) va -= S;
      const double vb = mir[a * TJ + TJ - 1 - b0 - e];
      split3d(va + vb, uu[0][e], uu[1][e], uu[2][e]);
      split3d(va - vb, ww[0][e], ww[1][e], ww[2][e]);
    }
    const long q = bofs(i, j, h);
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
      u32x2 pu, pw;
      pu.x = uu[pl][0] | (unsigned(uu[pl][1]) << 16); pu.y = uu[pl][2] | (unsigned(uu[pl][3]) << 16);
      pw.x = ww[pl][0] | (unsigned(ww[pl][1]) << 16); pw.y = ww[pl][2] | (unsigned(ww[pl][3]) << 16);
      *reinterpret_cast<OTM_GLOBAL u32x2*>(out + pl * nh + q) = pu;
      *reinterpret_cast<OTM_GLOBAL u32x2*>(out + (3 + pl) * nh + q) = pw;
    }
  }
}

// forward positions = origin point double(i)*(1/n) + vertex gradient of psi
// (otsolver_2dgrid.cpp:223-227); the final potential, in the buffer of the
// last completed iteration, goes back to its home buffer f.psi
__global__ __launch_bounds__(NT) void k_forward(JOB_ARGS) {
  const Loc L = locate_tile(starts, count);
  const Job& J = jobs[L.p];
  const int n = J.f.n;
  const auto st = gp((const SolveState*)J.f.state);
  const double* fin = J.f.psi_buf(st->it);
  const auto psi = gp(fin);
  const auto G = gp(J.fwd_out);
  __shared__ double gx[NV], gy[NV];
  __shared__ double s[NF];
  int i0, j0;
  tile_origin(L.t, n, i0, j0);
  load_halo(s, n, i0, j0, [&](long fc) { return psi[fc]; });
  __syncthreads();
  if (fin != J.f.psi) {
    const auto home = gp(J.f.psi);
    for (int t = threadIdx.x; t < TI * TJ; t += NT) {
      const int a = t / TJ, b = t % TJ;
      const int fi = i0 + a, fj = j0 + b;
      if (fi < n && fj < n) home[long(fi) * n + fj] = s[(a + 1) * LDF + b + 1];
    }
  }
  tile_grads(s, gx, gy, n, i0, j0);
  __syncthreads();
  const long V1 = n + 1;
  const long V = V1 * V1;
  const double dx = 1. / double(n);
  for (int t = threadIdx.x; t < NV; t += NT) {
    const int va = t / LDV, vb = t - va * LDV;
    const int vi = i0 + va, vj = j0 + vb;
    const bool own_i = (va < TI) || (vi == n);
    const bool own_j = (vb < TJ) || (vj == n);
    if (own_i && own_j && vi <= n && vj <= n) {
      const long v = vi * V1 + vj;
      G[v] = double(vi) * dx + gx[t];
      G[V + v] = double(vj) * dx + gy[t];
    }
  }
}

// number of jobs whose state is not done (one workgroup)
// out[0]: jobs still running; out[1]: those of them still far from their stop.
__global__ __launch_bounds__(NT) void k_count_open(const Job* __restrict__ jobs, int count, int* out) {
  __shared__ int total[2];
  if (threadIdx.x < 2) total[threadIdx.x] = 0;
  __syncthreads();
  int c = 0, far = 0;
  for (int p = threadIdx.x; p < count; p += NT) {
    const auto st = gp((const SolveState*)jobs[p].f.state);
    if (st->done) continue;
    ++c;
    far += far_from_stop(st->e, st->threshold, st->it, st->max_iter) ? 1 : 0;
  }
  if (c) atomicAdd(&total[0], c);
  if (far) atomicAdd(&total[1], far);
  __syncthreads();
  if (threadIdx.x < 2) gp(out)[threadIdx.x] = total[threadIdx.x];
}

__global__ __launch_bounds__(NT) void k_copy_adjusted(JOB_ARGS) {
  const Loc L = locate(starts, count);
  const Job& J = jobs[L.p];
  if (!J.adj_out) return;
  const auto out = gp(J.adj_out);
  const auto rho = gp(J.f.rho);
  for (long i = long(L.t) * NT + threadIdx.x; i < J.f.F; i += long(L.nb) * NT) out[i] = rho[i];
}

// vertex gradients of a face field into a (2V SoA) array; add_origin: forward
// positions = origin point double(i)*(1/n) + gradient (otsolver_2dgrid.cpp:223-227)
__global__ __launch_bounds__(NT) void k_vertex_gradients(const double* psi, double* G, int n, int add_origin) {
  __shared__ double gx[NV], gy[NV];
  __shared__ double s[NF];
  const int i0 = blockIdx.y * TI, j0 = blockIdx.x * TJ;
  load_halo(s, n, i0, j0, [&](long f) { return psi[f]; });
  __syncthreads();
  tile_grads(s, gx, gy, n, i0, j0);
  __syncthreads();
  const long V1 = n + 1;
  const long V = V1 * V1;
  const double dx = 1. / double(n);
  for (int t = threadIdx.x; t < NV; t += NT) {
    const int va = t / LDV, vb = t - va * LDV;
    const int vi = i0 + va, vj = j0 + vb;
    const bool own_i = (va < TI) || (vi == n);
    const bool own_j = (vb < TJ) || (vj == n);
    if (own_i && own_j && vi <= n && vj <= n) {
      const long v = vi * V1 + vj;
      if (add_origin) {
        G[v] = double(vi) * dx + gx[t];
        G[V + v] = double(vj) * dx + gy[t];
      } else {
        G[v] = gx[t];
        G[V + v] = gy[t];
      }
    }
  }
}

// residual only (component entry point): r = area(psi) - A rho, partial |r|^2
__global__ __launch_bounds__(NT) void k_residual_only(const double* psi, const double* rho, double* r, int n,
                                                     double* partials) {
  __shared__ double gx[NV], gy[NV];
  __shared__ double s[NF];
  const int i0 = blockIdx.y * TI, j0 = blockIdx.x * TJ;
  load_halo(s, n, i0, j0, [&](long f) { return psi[f]; });
  __syncthreads();
  tile_grads(s, gx, gy, n, i0, j0);
  __syncthreads();
  const double e = 1. / double(n), A = 1.0 / (double(n) * double(n));
  double acc[1] = {0.};
  for (int t = threadIdx.x; t < TI * TJ; t += NT) {
    const int a = t / TJ, b = t % TJ;
    const int fi = i0 + a, fj = j0 + b;
    if (fi < n && fj < n) {
      const long f = long(fi) * n + fj;
      const double rv = quad_area(gx, gy, a, b, e) - A * rho[f];
      r[f] = rv;
      acc[0] += rv * rv;
    }
  }
  write_partials<1>(acc, partials);
}

// quartic coefficients only (component entry point)
__global__ __launch_bounds__(NT) void k_params_1d(const double* g0, const double* dir, double* aout, double* bout,
                                                 int n) {
  __shared__ double gx[NV], gy[NV], hx[NV], hy[NV];
  __shared__ double s[NF];
  const int i0 = blockIdx.y * TI, j0 = blockIdx.x * TJ;
  load_halo(s, n, i0, j0, [&](long f) { return dir[f]; });
  const long V1 = n + 1;
  const long V = V1 * V1;
  for (int t = threadIdx.x; t < NV; t += NT) {
    const int a = t / LDV, b = t - a * LDV;
    const int vi = i0 + a, vj = j0 + b;
    const bool in = vi <= n && vj <= n;
    hx[t] = in ? g0[vi * V1 + vj] : 0.;
    hy[t] = in ? g0[V + vi * V1 + vj] : 0.;
  }
  __syncthreads();
  tile_grads(s, gx, gy, n, i0, j0);
  __syncthreads();
  const double e = 1. / double(n);
  for (int t = threadIdx.x; t < TI * TJ; t += NT) {
    const int a = t / TJ, b = t % TJ;
    const int fi = i0 + a, fj = j0 + b;
    if (fi < n && fj < n) {
      const long f = long(fi) * n + fj;
      const int v00 = a * LDV + b, v10 = (a + 1) * LDV + b, v01 = a * LDV + b + 1, v11 = (a + 1) * LDV + b + 1;
      const double d0ax = hx[v11] - hx[v00] + e;
      const double d0ay = hy[v11] - hy[v00] + e;
      const double d0bx = hx[v01] - hx[v10] - e;
      const double d0by = hy[v01] - hy[v10] + e;
      const double ddax = gx[v11] - gx[v00];
      const double dday = gy[v11] - gy[v00];
      const double ddbx = gx[v01] - gx[v10];
      const double ddby = gy[v01] - gy[v10];
      aout[f] = 0.5 * (ddax * ddby - dday * ddbx);
      bout[f] = 0.5 * (ddax * d0by - dday * d0bx + d0ax * ddby - d0ay * ddbx);
    }
  }
}

inline dim3 tile_grid(int n) { return dim3(tiles_j(n), tiles_i(n)); }

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_fill(double* p, long count, double v, hipStream_t s) {
  if (count <= 0) return;
  const long blocks = std::min<long>((count + NT - 1) / NT, 4096);
  hipLaunchKernelGGL(k_fill, dim3(unsigned(blocks)), dim3(NT), 0, s, p, count, v);
  OTM_HIP(hipGetLastError());
}

namespace {
#define JOB_PASS(kind) (const Job*)js.jobs, (const int*)js.starts_of(kind), js.count
inline dim3 grid_of(const JobSet& js, int kind) { return dim3(unsigned(js.total[kind])); }
}  // namespace

// density adjustment, field initialisation, initial residual and solver state
void launch_start(const JobSet& js, double max_ratio, int max_iter, double threshold, int cj, hipStream_t s,
                  int* flags) {
  hipLaunchKernelGGL(k_adjust_reduce, grid_of(js, KIND_FLAT), dim3(NT), 0, s, JOB_PASS(KIND_FLAT));
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_adjust_scalar, dim3(js.count), dim3(NT), 0, s, JOB_PASS(KIND_FLAT), max_ratio);
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_adjust_apply, grid_of(js, KIND_FLAT), dim3(NT), 0, s, JOB_PASS(KIND_FLAT));
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_init_fields, grid_of(js, KIND_FLAT), dim3(NT), 0, s, JOB_PASS(KIND_FLAT));
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_initial_residual, grid_of(js, KIND_STENCIL), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL));
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scalar_init, dim3(js.count), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL), max_iter, threshold, cj,
                     js.count == 1 ? flags : nullptr);
  OTM_HIP(hipGetLastError());
}

// Scalar kernels of the iteration in the wide form when a job has more
// stencil tiles than one fold batch (n > 1024); a job's result does not
// depend on the form as long as it has at most that many (and, of the
// transform, at most that many |d_hat|^2 partials, which holds up to 1024^2).
static bool wide_fold(const JobSet& js) {
  for (const Job& J : js.host)
    if (num_tiles(J.f.n) > NT * FOLD_BATCH) return true;
  return false;
}

// the |d_hat|^2 partials of each job were written by the workgroups of the
// last stage of its transform (KIND_DNORM)
void launch_beta(const JobSet& js, hipStream_t s, int* flags) {
  hipLaunchKernelGGL(k_beta, grid_of(js, KIND_STENCIL), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL));
  OTM_HIP(hipGetLastError());
  if (wide_fold(js))
    hipLaunchKernelGGL(k_scalar_beta<true>, dim3(js.count), dim3(BETA_GROUPS * NT), 0, s, JOB_PASS(KIND_STENCIL),
                       (const int*)js.starts_of(KIND_DNORM), js.count == 1 ? flags : nullptr);
  else
    hipLaunchKernelGGL(k_scalar_beta<false>, dim3(js.count), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL),
                       (const int*)js.starts_of(KIND_DNORM), js.count == 1 ? flags : nullptr);
  OTM_HIP(hipGetLastError());
}

void launch_dir_coef_pass(const JobSet& js, hipStream_t s) {
  hipLaunchKernelGGL(k_dir_coef, grid_of(js, KIND_STENCIL), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL));
  OTM_HIP(hipGetLastError());
}

void launch_dir_coef(const JobSet& js, hipStream_t s, int* flags) {
  launch_dir_coef_pass(js, s);
  if (wide_fold(js))
    hipLaunchKernelGGL(k_scalar_step<true>, dim3(js.count), dim3(STEP_GROUPS * NT), 0, s, JOB_PASS(KIND_STENCIL),
                       js.count == 1 ? flags : nullptr);
  else
    hipLaunchKernelGGL(k_scalar_step<false>, dim3(js.count), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL),
                       js.count == 1 ? flags : nullptr);
  OTM_HIP(hipGetLastError());
}

// jobs on the FP64 transform: plain update; jobs on the mixed transform: the
// update that also folds and splits their next right-hand side
void launch_update(const JobSet& js, hipStream_t s) {
  if (js.total[KIND_STENCIL64] > 0) {
    hipLaunchKernelGGL(k_update, grid_of(js, KIND_STENCIL64), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL64));
    OTM_HIP(hipGetLastError());
  }
  if (js.total[KIND_UPAIR] > 0) {
    hipLaunchKernelGGL(k_update_fold, grid_of(js, KIND_UPAIR), dim3(NT), 0, s, JOB_PASS(KIND_UPAIR));
    OTM_HIP(hipGetLastError());
  }
}

void launch_finish(const JobSet& js, hipStream_t s) {
  hipLaunchKernelGGL(k_forward, grid_of(js, KIND_STENCIL), dim3(NT), 0, s, JOB_PASS(KIND_STENCIL));
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_copy_adjusted, grid_of(js, KIND_FLAT), dim3(NT), 0, s, JOB_PASS(KIND_FLAT));
  OTM_HIP(hipGetLastError());
}

void launch_count_open(const JobSet& js, int* d_open, hipStream_t s) {
  hipLaunchKernelGGL(k_count_open, dim3(1), dim3(NT), 0, s, (const Job*)js.jobs, js.count, d_open);
  OTM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// job tables
// ---------------------------------------------------------------------------
namespace {

// Device image of the launch kinds for the table's jobs and their assigned
// tile configurations (every valid configuration in a tuning table): per
// kind, count + 1 workgroup prefix sums, then the job of each workgroup.
// Fills js.total and js.kofs.
std::vector<int> job_starts(JobSet& js) {
  js.total.assign(KIND_COUNT, 0);
  js.kofs.assign(KIND_COUNT, 0);
  std::vector<int> img;
  // GEMM launches start every job on a multiple of 8 workgroups (XCD-aware tile order)
  auto pad8 = [](int w) { return (w + 7) / 8 * 8; };
  std::vector<int> w(js.count);
  for (int kind = 0; kind < KIND_COUNT; ++kind) {
    for (int p = 0; p < js.count; ++p) {
      const Job& J = js.host[p];
      const int n = J.f.n;
      const auto& fc = js.jcfg[p];
      const auto& mc = js.jmcfg[p];
      // an autotuner table covers every job with both GEMM transforms
      const bool r64 = js.tuning || js.jroute[p] == 0, rmx = js.tuning || js.jroute[p] == 1;
      const bool rfft = !js.tuning && js.jroute[p] == 2;
      int v = 0;
      if (kind == KIND_STENCIL) v = num_tiles(n);
      else if (kind == KIND_FLAT) v = flat_blocks(J.f.F);
      else if (kind == KIND_STENCIL64) v = (r64 || rfft) ? num_tiles(n) : 0;  // the plain update pass
      else if (kind == KIND_FFT) v = rfft && fft_pack(n) ? fft_pack_wgs(n) : 0;
      else if (kind == KIND_FFTW8) v = rfft && n == 1024 ? n / FFT_ROW_W : 0;
      else if (kind == KIND_FFTW4) v = rfft && n == 512 ? n / FFT_ROW_W : 0;
      else if (kind == KIND_FFTW16) v = rfft && n == 2048 ? n / FFT_ROW_W : 0;
      else if (kind == KIND_FFT2) v = rfft && fft_rows(n) == 4 && !fft_row_form(n) ? n / 4 : 0;
      else if (kind == KIND_FLATMX) v = rmx && mixed_supported(n) ? flat_blocks(J.f.F) : 0;
      else if (kind == KIND_PLAIN64) v = r64 ? ((n + 63) / 64) * ((n + 63) / 64) : 0;
      else if (kind == KIND_MRES) v = rmx && mixed_supported(n) ? mres_tiles(n) : 0;
      else if (kind == KIND_UPAIR) v = rmx && mixed_supported(n) ? upair_tiles(n) : 0;
      else if (kind == KIND_DNORM) {
        if (js.tuning) v = 0;
        else if (js.jroute[p] == 2) v = fft_wgs(n);
        else if (js.jroute[p] == 1) v = mixed_cfg_valid(3, n, mc[3]) ? pad8(mixed_tiles(3, n, mc[3])) : 0;
        else if (js.fold) v = fold_cfg_valid(n, fc[3]) ? pad8(fold_tiles(n, 3, fc[3])) : 0;
        else v = ((n + 63) / 64) * ((n + 63) / 64);
      } else if (kind < KIND_GEMM0) {
        const int g = kind - KIND_M0, stage = g / NMIXED, c = g % NMIXED;
        const bool on = rmx && mixed_cfg_valid(stage, n, c) && (js.tuning || mc[stage] == c);
        v = on ? pad8(mixed_tiles(stage, n, c)) : 0;
      } else {
        const int g = kind - KIND_GEMM0, stage = g / NFOLD, c = g % NFOLD;
        const bool on = r64 && fold_cfg_valid(n, c) && (js.tuning || fc[stage] == c);
        v = on ? pad8(fold_tiles(n, stage, c)) : 0;
      }
      w[p] = v;
    }
    js.kofs[kind] = long(img.size());
    int acc = 0;
    for (int p = 0; p < js.count; ++p) {
      img.push_back(acc);
      acc += w[p];
    }
    img.push_back(acc);
    js.total[kind] = acc;
    for (int p = 0; p < js.count; ++p) img.insert(img.end(), size_t(w[p]), p);
  }
  return img;
}

// a process-wide unique stamp per table assignment
long next_version() {
  static std::atomic<long> counter{0};
  return ++counter;
}

// configurations of the largest grid, reported by the configuration queries
void report_configs(JobSet& js) {
  int big = 0;
  for (int p = 1; p < js.count; ++p)
    if (js.host[p].f.n > js.host[big].f.n) big = p;
  for (int k = 0; k < 4; ++k) {
    js.cfg[k] = js.jcfg[big][k];
    js.mcfg[k] = js.jmcfg[big][k];
  }
}

}  // namespace

void jobset_create(JobSet& js, const std::vector<Job>& jobs, bool tuning) {
  js.count = int(jobs.size());
  js.host = jobs;
  js.tuning = tuning;
  js.fold = true;
  for (const Job& J : jobs) js.fold = js.fold && J.P != nullptr;
  js.jcfg.assign(js.count, {FOLD_DEFAULT, FOLD_DEFAULT, FOLD_DEFAULT, FOLD_DEFAULT});
  js.jmcfg.assign(js.count, {MIXED_DEFAULT, MIXED_DEFAULT, MIXED_DEFAULT, MIXED_DEFAULT});
  js.jroute.assign(js.count, 0);
  js.routed_mixed = 0;
  js.routed_fft = 0;
  report_configs(js);
  js.version = next_version();
  const std::vector<int> st = job_starts(js);
  OTM_HIP(hipMalloc(&js.jobs, js.count * sizeof(Job)));
  OTM_HIP(hipMalloc(&js.starts, st.size() * sizeof(int)));
  OTM_HIP(hipMemcpy(js.jobs, jobs.data(), js.count * sizeof(Job), hipMemcpyHostToDevice));
  OTM_HIP(hipMemcpy(js.starts, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice));
}

void jobset_assign(JobSet& js) {
  // no launch of the device may still read the old table image
  OTM_HIP(hipDeviceSynchronize());
  report_configs(js);
  const std::vector<int> st = job_starts(js);
  if (js.starts) (void)hipFree(js.starts);
  js.starts = nullptr;
  OTM_HIP(hipMalloc(&js.starts, st.size() * sizeof(int)));
  OTM_HIP(hipMemcpy(js.starts, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice));
  js.version = next_version();
}

void jobset_route(JobSet& js, const std::vector<int>& route) {
  if (route == js.jroute) return;
  js.jroute = route;
  js.routed_mixed = 0;
  js.routed_fft = 0;
  for (int r : route) {
    js.routed_mixed += r == 1 ? 1 : 0;
    js.routed_fft += r == 2 ? 1 : 0;
  }
  jobset_assign(js);
}

void jobset_upload(JobSet& js, hipStream_t s) {
  OTM_HIP(hipMemcpyAsync(js.jobs, js.host.data(), js.count * sizeof(Job), hipMemcpyHostToDevice, s));
}

void jobset_destroy(JobSet& js) {
  if (js.jobs) (void)hipFree(js.jobs);
  if (js.starts) (void)hipFree(js.starts);
  js = JobSet();
}

// step choice of the line search for `count` coefficient vectors z (5 each):
// the same pick_step the scalar kernel of a solve runs
__global__ __launch_bounds__(64) void k_pick_steps(const double* __restrict__ z, int count, double* __restrict__ alpha,
                                                  double* __restrict__ emin) {
  const int k = blockIdx.x * 64 + threadIdx.x;
  if (k >= count) return;
  double zz[5];
  for (int i = 0; i < 5; ++i) zz[i] = z[5 * long(k) + i];
  double e = 0.;
  alpha[k] = pick_step(zz, e);
  emin[k] = e;
}

void launch_pick_steps(const double* z, int count, double* alpha, double* emin, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_pick_steps, dim3((count + 63) / 64), dim3(64), 0, s, z, count, alpha, emin);
  OTM_HIP(hipGetLastError());
}

void launch_vertex_gradients(const double* psi, double* G, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_vertex_gradients, tile_grid(n), dim3(NT), 0, s, psi, G, n, 0);
  OTM_HIP(hipGetLastError());
}

void launch_residual_only(const double* psi, const double* rho, double* r, double* partials, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_residual_only, tile_grid(n), dim3(NT), 0, s, psi, rho, r, n, partials);
  OTM_HIP(hipGetLastError());
}

void launch_params_1d(const double* g0, const double* dir, double* a, double* b, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_params_1d, tile_grid(n), dim3(NT), 0, s, g0, dir, a, b, n);
  OTM_HIP(hipGetLastError());
}

}  // namespace otm
