// ============================================================================
// otmap oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// An FP64 CPU restatement of the reference's grid transport solver
// (ggael/otmap), written from the reference's behaviour, used only as the
// checker by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
// leg.  The product path (otmap_amd, libotmap_amd.so) never links or calls it.
//
// Each function cites the reference lines it restates (paths relative to the
// reference root).  Arithmetic is kept in the reference's operation order
// where the reference fixes it (stencils, predicates, BVH split), and the file
// is compiled with -ffp-contract=off so no FMA contraction changes rounding.
//
// Pinning (how each function is checked independently of this file):
//   adjust_density        closed form (integral 1, ratio cap) + reference goldens
//   vertex_gradients      hand-derived n=2,3 stencils + reference goldens
//   face_area/residual    invariants (sum area = 1, psi=const -> area = A) + goldens
//   laplacian (LDL^T)     dense solve / scipy DCT-II closed form + goldens
//   1d params / solve_1d  exactness of r(psi+t d) = r + b t + a t^2 at 3 t's,
//                         brute-force scan of e(t) + goldens
//   solve (driver)        separable-density closed form, constant density,
//                         reference iteration/alpha/beta traces
//   BVH / inverse         brute-force all-faces location + reference goldens
//   rasterize             brute-force per-pixel scan + reference goldens
//   transport costs       reference goldens
// ============================================================================

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

inline int fidx(int n, int i, int j) { return j + i * n; }
inline int vidx(int n, int i, int j) { return j + i * (n + 1); }

// ---------------------------------------------------------------------------
// Density adjustment  (otlib/otsolver_2dgrid.cpp:47-73)
// ---------------------------------------------------------------------------
void adjust_density(int n, double* rho, double max_ratio) {
  const long F = long(n) * n;
  const double A = 1.0 / (double(n) * double(n));
  double s = 0;
  for (long f = 0; f < F; ++f) s += rho[f];
  const double I = s * A;
  for (long f = 0; f < F; ++f) rho[f] /= I;
  double mn = rho[0], mx = rho[0];
  for (long f = 1; f < F; ++f) { mn = std::min(mn, rho[f]); mx = std::max(mx, rho[f]); }
  const double ratio = std::min(max_ratio, mx / mn);
  const double cand = (ratio * mn - mx) / (1. - ratio);
  const double beta = (0. < cand) ? cand : 0.;  // std::max(0., cand): NaN -> 0
  for (long f = 0; f < F; ++f) rho[f] = (rho[f] + beta) / (1. + beta);
}

// ---------------------------------------------------------------------------
// Vertex gradients of a face potential  (otlib/otsolver_2dgrid.cpp:362-423)
// G is SoA: Gx[V] then Gy[V].  Interior columns j in [1, s) follow the 2-wide
// SIMD packet order (p10+p11)-(p00+p01); columns j in [s, n) follow the scalar
// tail order ((p10+p11)-p00)-p01, with s = 2*floor((n-1)/2) (SSE2 packets).
// ---------------------------------------------------------------------------
void vertex_gradients(int n, const double* psi, double* G) {
  const long V = long(n + 1) * (n + 1);
  double* Gx = G;
  double* Gy = G + V;
  const double w = double(n);
  const double hw = 0.5 * w;
  const int s = ((n - 1) / 2) * 2;
  for (int i = 1; i < n; ++i) {
    const double* r0 = psi + fidx(n, i - 1, 0);
    const double* r1 = psi + fidx(n, i, 0);
    for (int j = 1; j < n; ++j) {
      const double p00 = r0[j - 1], p01 = r0[j], p10 = r1[j - 1], p11 = r1[j];
      const int v = vidx(n, i, j);
      if (j < s) {
        Gx[v] = hw * ((p10 + p11) - (p00 + p01));
        Gy[v] = hw * ((p01 + p11) - (p00 + p10));
      } else {
        Gx[v] = hw * (((p10 + p11) - p00) - p01);
        Gy[v] = hw * (((p01 + p11) - p00) - p10);
      }
    }
  }
  for (int k = 1; k < n; ++k) {
    Gx[vidx(n, k, 0)] = w * (psi[fidx(n, k, 0)] - psi[fidx(n, k - 1, 0)]);
    Gy[vidx(n, k, 0)] = 0.;
    Gx[vidx(n, k, n)] = w * (psi[fidx(n, k, n - 1)] - psi[fidx(n, k - 1, n - 1)]);
    Gy[vidx(n, k, n)] = 0.;
    Gx[vidx(n, 0, k)] = 0.;
    Gy[vidx(n, 0, k)] = w * (psi[fidx(n, 0, k)] - psi[fidx(n, 0, k - 1)]);
    Gx[vidx(n, n, k)] = 0.;
    Gy[vidx(n, n, k)] = w * (psi[fidx(n, n - 1, k)] - psi[fidx(n, n - 1, k - 1)]);
  }
  const int corners[4] = {vidx(n, 0, 0), vidx(n, n, 0), vidx(n, 0, n), vidx(n, n, n)};
  for (int c : corners) { Gx[c] = 0.; Gy[c] = 0.; }
}

// ---------------------------------------------------------------------------
// Quad areas from vertex displacements  (otlib/otsolver_2dgrid.cpp:425-447)
// ---------------------------------------------------------------------------
// Precision experiment (not the reference method): area_mode 1 evaluates
// every quad area of the residual in FP32 arithmetic from the FP64 vertex
// gradients rounded to FP32 (e in FP32 too), the result widened back to FP64;
// 0 (default) is the reference's FP64.  Set by orc_set_area_precision.
int g_area_mode = 0;

void face_area(int n, const double* G, double* area) {
  const long V = long(n + 1) * (n + 1);
  const double* Gx = G;
  const double* Gy = G + V;
  const double e = 1. / double(n);
  if (g_area_mode == 1) {
    const float ef = float(e);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const int v00 = vidx(n, i, j), v10 = vidx(n, i + 1, j), v01 = vidx(n, i, j + 1), v11 = vidx(n, i + 1, j + 1);
        const float a = float(Gx[v11]) - float(Gx[v00]) + ef, b = float(Gy[v01]) - float(Gy[v10]) + ef;
        const float c = float(Gy[v11]) - float(Gy[v00]) + ef, d = float(Gx[v01]) - float(Gx[v10]) - ef;
        area[fidx(n, i, j)] = double(0.5f * (a * b - c * d));
      }
    return;
  }
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      const int v00 = vidx(n, i, j), v10 = vidx(n, i + 1, j), v01 = vidx(n, i, j + 1), v11 = vidx(n, i + 1, j + 1);
      area[fidx(n, i, j)] = 0.5 * ((Gx[v11] - Gx[v00] + e) * (Gy[v01] - Gy[v10] + e) -
                                   (Gy[v11] - Gy[v00] + e) * (Gx[v01] - Gx[v10] - e));
    }
  }
}

// residual r = area(psi) - A*rho, returns |r|^2/A  (otlib/otsolver_2dgrid.cpp:449-465)
double residual(int n, const double* rho, const double* psi, double* r, std::vector<double>& G) {
  const long F = long(n) * n, V = long(n + 1) * (n + 1);
  const double A = 1.0 / (double(n) * double(n));
  G.resize(2 * V);
  vertex_gradients(n, psi, G.data());
  face_area(n, G.data(), r);
  double s = 0;
  for (long f = 0; f < F; ++f) { r[f] = r[f] - A * rho[f]; s += r[f] * r[f]; }
  return s / A;
}

// ---------------------------------------------------------------------------
// Nested-dissection ordering  (otlib/details/nested_dissection.cpp:14-61)
// Blocks of <= 4 rows or cols are emitted row by row; larger blocks emit the
// middle column then the middle row (minus the crossing) and recurse on the
// four quadrants; the whole list is reversed so that separators come last.
// ---------------------------------------------------------------------------
void nd_block(int i0, int j0, int rows, int cols, int stride, std::vector<int>& out) {
  const int i1 = i0 + rows, j1 = j0 + cols;
  if (rows <= 4 || cols <= 4) {
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j) out.push_back(i + j * stride);
    return;
  }
  const int si = i0 + rows / 2, sj = j0 + cols / 2;
  for (int i = i0; i < i1; ++i) out.push_back(i + stride * sj);
  for (int j = j0; j < j1; ++j)
    if (j != sj) out.push_back(si + stride * j);
  nd_block(i0, j0, si - i0, sj - j0, stride, out);
  nd_block(si + 1, j0, i1 - si - 1, sj - j0, stride, out);
  nd_block(si + 1, sj + 1, i1 - si - 1, j1 - sj - 1, stride, out);
  nd_block(i0, sj + 1, si - i0, j1 - sj - 1, stride, out);
}

std::vector<int> nd_ordering(int n) {
  std::vector<int> p;
  p.reserve(size_t(n) * n);
  nd_block(0, 0, n, n, n, p);
  std::reverse(p.begin(), p.end());
  return p;
}

// ---------------------------------------------------------------------------
// Pinned pseudo-Laplacian and its sparse LDL^T factor
// (assembly otlib/otsolver_2dgrid.cpp:284-321; factor/solve :336-340, :175-179).
// The factorisation is the up-looking LDL^T of T. Davis (ACM TOMS 31(4), 2005),
// restated: elimination tree + column counts, then row k of L from a sparse
// triangular solve over the etree reach of column k.
// ---------------------------------------------------------------------------
struct Laplacian {
  int n = 0;
  long N = 0;
  std::vector<int> perm, pinv;             // new k <- old perm[k]
  std::vector<long> Lp;
  std::vector<int> Li, parent;
  std::vector<double> Lx, D;
  bool ok = false;
};

Laplacian* laplacian_create(int n) {
  auto* L = new Laplacian;
  L->n = n;
  const long N = long(n) * n;
  L->N = N;
  // Row-wise assembly of M = -L: weight -1/2 at the 4 clamped diagonal
  // neighbours, diagonal = -(sum of weights) = 2; duplicates are summed.
  std::vector<std::vector<std::pair<int, double>>> rows(N);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      const int id = fidx(n, i, j);
      const int r1 = i == 0 ? i : i - 1, c1 = j == 0 ? j : j - 1;
      const int r2 = i == n - 1 ? i : i + 1, c2 = j == n - 1 ? j : j + 1;
      const int nb[4] = {fidx(n, r1, c1), fidx(n, r1, c2), fidx(n, r2, c1), fidx(n, r2, c2)};
      auto add = [&](int c, double v) {
        for (auto& e : rows[id])
          if (e.first == c) { e.second += v; return; }
        rows[id].push_back({c, v});
      };
      double sw = 0;
      for (int k = 0; k < 4; ++k) { add(nb[k], -0.5); sw += -0.5; }
      add(id, -sw);
    }
  }
  // weakly enforce psi(0,0) = 0
  for (auto& e : rows[0])
    if (e.first == 0) e.second += std::abs(e.second) * 1e4;

  L->perm = nd_ordering(n);
  L->pinv.assign(N, 0);
  for (long k = 0; k < N; ++k) L->pinv[L->perm[k]] = int(k);

  // upper triangle of C = P M P^T, column-compressed
  std::vector<std::vector<std::pair<int, double>>> cols(N);
  for (long r = 0; r < N; ++r)
    for (auto& e : rows[r]) {
      const int pi = L->pinv[r], pj = L->pinv[e.first];
      if (pi <= pj) cols[pj].push_back({pi, e.second});
    }

  // symbolic: elimination tree and column counts
  std::vector<int>& parent = L->parent;
  parent.assign(N, -1);
  std::vector<int> flag(N, -1);
  std::vector<long> Lnz(N, 0);
  for (long k = 0; k < N; ++k) {
    flag[k] = int(k);
    for (auto& e : cols[k]) {
      int i = e.first;
      for (; flag[i] != k; i = parent[i]) {
        if (parent[i] == -1) parent[i] = int(k);
        ++Lnz[i];
        flag[i] = int(k);
      }
    }
  }
  L->Lp.assign(N + 1, 0);
  for (long k = 0; k < N; ++k) L->Lp[k + 1] = L->Lp[k] + Lnz[k];
  const long nnz = L->Lp[N];
  L->Li.assign(nnz, 0);
  L->Lx.assign(nnz, 0.);
  L->D.assign(N, 0.);

  // numeric
  std::vector<double> Y(N, 0.);
  std::vector<int> pattern(N);
  std::fill(Lnz.begin(), Lnz.end(), 0);
  L->ok = true;
  for (long k = 0; k < N; ++k) {
    long top = N;
    flag[k] = int(k);
    for (auto& e : cols[k]) {
      int i = e.first;
      Y[i] += e.second;
      long len = 0;
      for (; flag[i] != k; i = parent[i]) { pattern[len++] = i; flag[i] = int(k); }
      while (len > 0) pattern[--top] = pattern[--len];
    }
    double dk = Y[k];
    Y[k] = 0.;
    for (; top < N; ++top) {
      const int i = pattern[top];
      const double yi = Y[i];
      Y[i] = 0.;
      const long p0 = L->Lp[i], p1 = L->Lp[i] + Lnz[i];
      for (long p = p0; p < p1; ++p) Y[L->Li[p]] -= L->Lx[p] * yi;
      const double lki = yi / L->D[i];
      dk -= lki * yi;
      L->Li[p1] = int(k);
      L->Lx[p1] = lki;
      ++Lnz[i];
    }
    if (dk == 0.) L->ok = false;
    L->D[k] = dk;
  }
  return L;
}

// d = M^{-1} r, then d -= mean(d)   (otlib/otsolver_2dgrid.cpp:175-179)
void laplacian_solve(const Laplacian* L, const double* r, double* d) {
  const long N = L->N;
  std::vector<double> x(N);
  for (long k = 0; k < N; ++k) x[k] = r[L->perm[k]];
  for (long j = 0; j < N; ++j) {
    const double xj = x[j];
    for (long p = L->Lp[j]; p < L->Lp[j + 1]; ++p) x[L->Li[p]] -= L->Lx[p] * xj;
  }
  for (long j = 0; j < N; ++j) x[j] /= L->D[j];
  for (long j = N - 1; j >= 0; --j) {
    double xj = x[j];
    for (long p = L->Lp[j]; p < L->Lp[j + 1]; ++p) xj -= L->Lx[p] * x[L->Li[p]];
    x[j] = xj;
  }
  double mean = 0;
  for (long k = 0; k < N; ++k) { d[L->perm[k]] = x[k]; }
  for (long k = 0; k < N; ++k) mean += d[k];
  mean /= double(N);
  for (long k = 0; k < N; ++k) d[k] -= mean;
}

// ---------------------------------------------------------------------------
// Quartic line-search coefficients  (otlib/otsolver_2dgrid.cpp:467-539)
// r(psi + t*dir) = r(psi) + b*t + a*t^2 exactly, per quad.
// ---------------------------------------------------------------------------
void params_1d(int n, const double* g0, const double* gd, double* a, double* b) {
  const long V = long(n + 1) * (n + 1);
  const double e = 1. / double(n);
  const double *g0x = g0, *g0y = g0 + V, *gdx = gd, *gdy = gd + V;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      const int v00 = vidx(n, i, j), v10 = vidx(n, i + 1, j), v01 = vidx(n, i, j + 1), v11 = vidx(n, i + 1, j + 1);
      const double d0ax = g0x[v11] - g0x[v00] + e;
      const double d0ay = g0y[v11] - g0y[v00] + e;
      const double d0bx = g0x[v01] - g0x[v10] - e;
      const double d0by = g0y[v01] - g0y[v10] + e;
      const double ddax = gdx[v11] - gdx[v00];
      const double dday = gdy[v11] - gdy[v00];
      const double ddbx = gdx[v01] - gdx[v10];
      const double ddby = gdy[v01] - gdy[v10];
      const int id = fidx(n, i, j);
      a[id] = 0.5 * (ddax * ddby - dday * ddbx);
      b[id] = 0.5 * (ddax * d0by - dday * d0bx + d0ax * ddby - d0ay * ddbx);
    }
  }
}

// Real roots of the monic cubic t^3 + c2 t^2 + c1 t + c0 by bracketing each
// monotone piece (split at the critical points) and bisecting to the last bit.
int cubic_real_roots(double c2, double c1, double c0, double* roots) {
  auto f = [&](double t) { return ((t + c2) * t + c1) * t + c0; };
  auto bisect = [&](double lo, double hi) {
    double flo = f(lo);
    for (int it = 0; it < 2000; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      const double fm = f(mid);
      if (fm == 0.) return mid;
      if ((fm < 0.) == (flo < 0.)) { lo = mid; flo = fm; } else { hi = mid; }
    }
    return std::abs(f(lo)) <= std::abs(f(hi)) ? lo : hi;
  };
  // expand from x0 in direction dir until f changes sign w.r.t. f(x0)
  auto outward = [&](double x0, double dir) {
    const double f0 = f(x0);
    double step = 1.0 + std::abs(x0);
    double x1 = x0 + dir * step;
    for (int it = 0; it < 4000 && ((f(x1) < 0.) == (f0 < 0.)); ++it) {
      step *= 2.;
      x1 = x0 + dir * step;
    }
    return dir > 0 ? bisect(x0, x1) : bisect(x1, x0);
  };
  const double disc = c2 * c2 - 3. * c1;  // critical points of the cubic
  if (!(disc > 0.)) {
    roots[0] = f(0.) == 0. ? 0. : outward(0., f(0.) > 0. ? -1. : 1.);
    return 1;
  }
  const double sq = std::sqrt(disc);
  double t1 = (-c2 - sq) / 3., t2 = (-c2 + sq) / 3.;  // local max at t1, local min at t2
  const double f1 = f(t1), f2 = f(t2);
  int k = 0;
  if (f1 < 0.) {  // only the rightmost branch crosses zero
    roots[k++] = outward(t2, 1.);
  } else if (f2 > 0.) {  // only the leftmost branch
    roots[k++] = outward(t1, -1.);
  } else {
    roots[k++] = (f1 == 0.) ? t1 : outward(t1, -1.);
    if (f1 != 0. && f2 != 0.) roots[k++] = bisect(t1, t2);
    roots[k++] = (f2 == 0.) ? t2 : outward(t2, 1.);
  }
  return k;
}

// Exact quartic line search  (otlib/otsolver_2dgrid.cpp:541-606).
struct LineSearch {
  double z[5];
  double alpha;
  double emin;
  int nroots;
};

LineSearch pick_step(const double z[5]) {
  LineSearch ls;
  for (int i = 0; i < 5; ++i) ls.z[i] = z[i];
  const double w0 = z[1], w1 = 2. * z[2], w2 = 3. * z[3], w3 = 4. * z[4];
  const double c0 = w0 / w3, c1 = w1 / w3, c2 = w2 / w3;
  double roots[3];
  int nr = 0;
  if (std::isfinite(c0) && std::isfinite(c1) && std::isfinite(c2)) nr = cubic_real_roots(c2, c1, c0, roots);
  ls.nroots = nr;
  bool found = false;
  double rmin = 0, alpha = 0;
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
  ls.alpha = alpha;
  ls.emin = rmin;
  return ls;
}

double dot(const double* x, const double* y, long N) {
  double s = 0;
  for (long k = 0; k < N; ++k) s += x[k] * y[k];
  return s;
}

// ---------------------------------------------------------------------------
// Grid mesh geometry and predicates
// (otlib/utils/mesh_utils.cpp:18-88, mesh_utils.h:18-25, eigen_addons.h:16-23)
// ---------------------------------------------------------------------------
struct P2 { double x, y; };
inline double cross2(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }
inline P2 sub(P2 a, P2 b) { return {a.x - b.x, a.y - b.y}; }
inline double sarea(P2 a, P2 b, P2 c) { return 0.5 * cross2(sub(b, a), sub(c, a)); }

inline void tri_coords(P2 q, P2 p0, P2 p1, P2 p2, double& u, double& v) {
  const P2 q2 = sub(q, p2), eu = sub(p1, p2), ev = sub(p0, p2);
  const double a2 = cross2(ev, eu);
  u = cross2(q2, eu) / a2;
  v = cross2(ev, q2) / a2;
}

bool inside_quad(P2 q, const P2* p) {
  const double eps = 1e-8;
  double u, v;
  tri_coords(q, p[0], p[1], p[2], u, v);
  if (u >= -eps && v >= -eps && u + v <= 1. + eps) return true;
  tri_coords(q, p[0], p[2], p[3], u, v);
  if (u >= -eps && v >= -eps && u + v <= 1. + eps) return true;
  return false;
}

bool bilinear_uv(P2 q, const P2* p, double& u, double& v) {
  const double A0 = sarea(q, p[0], p[1]);
  const double A1 = sarea(q, p[1], p[2]);
  const double A2 = sarea(q, p[2], p[3]);
  const double A3 = sarea(q, p[3], p[0]);
  const double B0 = sarea(q, p[3], p[1]);
  const double B1 = sarea(q, p[0], p[2]);
  const double B3 = sarea(q, p[2], p[0]);
  const double D = B0 * B0 + B1 * B1 + 2 * A0 * A2 + 2 * A1 * A3;
  if (D < 0) return false;
  const double sD = std::sqrt(D);
  const double E0 = 2 * A0 - B0 - B1 + sD;
  const double E3 = 2 * A3 - B3 - B0 + sD;
  const double zero = 4 * DBL_MIN;
  v = std::abs(A0) <= zero ? 0 : 2 * A0 / E0;
  u = std::abs(A3) <= zero ? 0 : 2 * A3 / E3;
  const double E0p = 2 * A0 - B0 - B1 - sD;
  const double E3p = 2 * A3 - B3 - B0 - sD;
  double v_alt = std::abs(A0) <= zero ? 0 : 2 * A0 / E0p;
  double u_alt = std::abs(A3) <= zero ? 0 : 2 * A3 / E3p;
  const double eps = std::sqrt(DBL_EPSILON);
  if (std::abs(u) < eps) u = 0;
  if (std::abs(v) < eps) v = 0;
  if (std::abs(u - 1.) < eps) u = 1;
  if (std::abs(v - 1.) < eps) v = 1;
  if (std::abs(u_alt) < eps) u_alt = 0;
  if (std::abs(v_alt) < eps) v_alt = 0;
  if (std::abs(u_alt - 1.) < eps) u_alt = 1;
  if (std::abs(v_alt - 1.) < eps) v_alt = 1;
  if ((u < 0 || v < 0 || u > 1 || v > 1) && (u_alt >= 0 && v_alt >= 0 && u_alt <= 1 && v_alt <= 1)) {
    std::swap(u, u_alt);
    std::swap(v, v_alt);
  }
  return true;
}

bool bilinear_weights(P2 q, const P2* p, double* w) {
  double u, v;
  if (!bilinear_uv(q, p, u, v)) return false;
  w[0] = (1. - u) * (1 - v);
  w[1] = u * (1. - v);
  w[2] = u * v;
  w[3] = (1. - u) * v;
  const double eps = 4 * DBL_EPSILON;
  for (int k = 0; k < 4; ++k) {
    if (std::abs(w[k]) <= eps) w[k] = 0;
    else if (std::abs(w[k] - 1.) <= eps) w[k] = 1;
  }
  return true;
}

// quad (i,j) -> vertices (i,j),(i+1,j),(i+1,j+1),(i,j+1)   (otlib/utils/mesh_utils.cpp:111-120)
inline void quad_vertices(int n, int f, int* v) {
  const int i = f / n, j = f % n;
  v[0] = vidx(n, i, j);
  v[1] = vidx(n, i + 1, j);
  v[2] = vidx(n, i + 1, j + 1);
  v[3] = vidx(n, i, j + 1);
}

// origin vertex position: double(i) * (1/n)   (otlib/utils/mesh_utils.cpp:99-108)
inline P2 origin_point(int n, int v) {
  const double dx = 1. / double(n);
  const int i = v / (n + 1), j = v % (n + 1);
  return {double(i) * dx, double(j) * dx};
}

// ---------------------------------------------------------------------------
// 2-D BVH over the quads of a deformed grid  (otlib/utils/bvh2d.cpp:29-243)
// ---------------------------------------------------------------------------
struct BNode {
  double lo[2], hi[2];
  int first;   // first child (inner) or first face slot (leaf)
  int count;   // faces in the leaf
  int leaf;
};

struct BVH {
  int n = 0;
  std::vector<P2> pts;
  std::vector<int> faces;
  std::vector<P2> cent;
  std::vector<BNode> nodes;
};

inline bool box_contains(const BNode& b, P2 p) {
  return b.lo[0] <= p.x && b.lo[1] <= p.y && p.x <= b.hi[0] && p.y <= b.hi[1];
}

int bvh_split(BVH& B, int start, int end, int dim, float sv) {
  auto c = [&](int k) { return dim == 0 ? B.cent[k].x : B.cent[k].y; };
  int l = start, r = end - 1;
  while (l < r) {
    while (l < end && c(l) < sv) ++l;
    while (r >= start && c(r) >= sv) --r;
    if (l >= r) break;
    std::swap(B.cent[l], B.cent[r]);
    std::swap(B.faces[l], B.faces[r]);
    ++l;
    --r;
  }
  if (l >= end) return end;  // both branches of the reference return end here
  return c(l) <= sv ? std::min(end, l + 1) : l;
}

void bvh_node(BVH& B, int id, int start, int end, int level, int leaf_size, int max_depth) {
  double lo[2] = {DBL_MAX, DBL_MAX}, hi[2] = {-DBL_MAX, -DBL_MAX};
  for (int k = start; k < end; ++k) {
    int v[4];
    quad_vertices(B.n, B.faces[k], v);
    for (int q = 0; q < 4; ++q) {
      const P2 p = B.pts[v[q]];
      lo[0] = std::min(lo[0], p.x); lo[1] = std::min(lo[1], p.y);
      hi[0] = std::max(hi[0], p.x); hi[1] = std::max(hi[1], p.y);
    }
  }
  const double eps = DBL_EPSILON;
  const double diag[2] = {hi[0] - lo[0], hi[1] - lo[1]};
  BNode nd;
  for (int d = 0; d < 2; ++d) {
    const double pad = diag[d] * (2. * eps) + eps;
    nd.lo[d] = lo[d] - pad;
    nd.hi[d] = hi[d] + pad;
  }
  nd.first = start;
  nd.count = std::max(0, end - start);
  nd.leaf = 1;
  B.nodes[id] = nd;
  if (end - start <= leaf_size || level >= max_depth) return;
  const int dim = diag[1] > diag[0] ? 1 : 0;
  const float sv = float(0.5 * (nd.hi[dim] + nd.lo[dim]));
  const int mid = bvh_split(B, start, end, dim, sv);
  if (mid == start || mid == end) return;
  const int child = int(B.nodes.size());
  B.nodes[id].leaf = 0;
  B.nodes[id].first = child;
  B.nodes[id].count = 0;
  B.nodes.resize(B.nodes.size() + 2);
  bvh_node(B, child, start, mid, level + 1, leaf_size, max_depth);
  bvh_node(B, child + 1, mid, end, level + 1, leaf_size, max_depth);
}

BVH* bvh_build(int n, const double* xy, int leaf_size, int max_depth) {
  auto* B = new BVH;
  B->n = n;
  const long V = long(n + 1) * (n + 1), F = long(n) * n;
  B->pts.resize(V);
  for (long v = 0; v < V; ++v) B->pts[v] = {xy[v], xy[V + v]};
  B->faces.resize(F);
  B->cent.resize(F);
  for (long f = 0; f < F; ++f) {
    int v[4];
    quad_vertices(n, int(f), v);
    P2 c = {0., 0.};
    for (int q = 0; q < 4; ++q) { c.x += B->pts[v[q]].x; c.y += B->pts[v[q]].y; }
    c.x /= 4; c.y /= 4;
    B->cent[f] = c;
    B->faces[f] = int(f);
  }
  B->nodes.resize(1);
  bvh_node(*B, 0, 0, int(F), 0, leaf_size, max_depth);
  return B;
}

struct Hit { int face; double w[4]; };

void bvh_visit(const BVH& B, int id, P2 q, std::vector<Hit>& hits, bool first_only) {
  const BNode& nd = B.nodes[id];
  if (nd.leaf) {
    for (int k = nd.first; k < nd.first + nd.count; ++k) {
      int v[4];
      quad_vertices(B.n, B.faces[k], v);
      P2 p[4] = {B.pts[v[0]], B.pts[v[1]], B.pts[v[2]], B.pts[v[3]]};
      if (!inside_quad(q, p)) continue;
      Hit h;
      if (!bilinear_weights(q, p, h.w)) continue;
      bool ok = true;
      for (int t = 0; t < 4; ++t) ok = ok && h.w[t] <= 1.0f && h.w[t] >= 0.0f;
      if (!ok) continue;
      h.face = B.faces[k];
      hits.push_back(h);
      if (first_only) return;
    }
    return;
  }
  for (int c = 0; c < 2; ++c) {
    if (box_contains(B.nodes[nd.first + c], q)) {
      bvh_visit(B, nd.first + c, q, hits, first_only);
      if (first_only && !hits.empty()) return;
    }
  }
}

void bvh_query(const BVH& B, P2 q, std::vector<Hit>& hits, bool first_only) {
  hits.clear();
  if (box_contains(B.nodes[0], q)) bvh_visit(B, 0, q, hits, first_only);
}

// TransportMap::inv_impl  (otlib/transport_map.cpp:50-102)
// Returns the number of hits; out = NaN when there is none.
int inverse_point(const BVH& B, const double* density, P2 pin, bool fast, double* out, int* face,
                  std::vector<Hit>& hits) {
  const P2 p = {std::min(std::max(pin.x, 0.), 1.), std::min(std::max(pin.y, 0.), 1.)};
  bvh_query(B, p, hits, fast);
  if (hits.empty()) {
    out[0] = out[1] = std::numeric_limits<double>::quiet_NaN();
    *face = -1;
    return 0;
  }
  int f = hits[0].face;
  const double* w = hits[0].w;
  if (!fast && hits.size() > 1) {
    double best = density[f];
    for (size_t k = 1; k < hits.size(); ++k) {
      if (density[hits[k].face] > best) {
        f = hits[k].face;
        w = hits[0].w;  // the reference keeps the first hit's weights
        best = density[hits[k].face] > best;  // ... and stores a bool
      }
    }
  }
  int v[4];
  quad_vertices(B.n, f, v);
  P2 r = origin_point(B.n, v[0]);
  double rx = w[0] * r.x, ry = w[0] * r.y;
  for (int k = 1; k < 4; ++k) {
    const P2 o = origin_point(B.n, v[k]);
    rx += w[k] * o.x;
    ry += w[k] * o.y;
  }
  out[0] = rx;
  out[1] = ry;
  *face = f;
  return int(hits.size());
}

}  // namespace

// ============================================================================
// C ABI of the oracle (loaded with ctypes by the tests and bench.py)
// ============================================================================
extern "C" {

void orc_adjust_density(int n, double* rho, double max_ratio) { adjust_density(n, rho, max_ratio); }

void orc_vertex_gradients(int n, const double* psi, double* G) { vertex_gradients(n, psi, G); }

void orc_face_area(int n, const double* G, double* area) { face_area(n, G, area); }
void orc_set_area_precision(int mode) { g_area_mode = mode; }

double orc_residual(int n, const double* rho, const double* psi, double* r) {
  std::vector<double> G;
  return residual(n, rho, psi, r, G);
}

void orc_nd_ordering(int n, int* perm) {
  auto p = nd_ordering(n);
  std::copy(p.begin(), p.end(), perm);
}

void* orc_laplacian_create(int n) { return laplacian_create(n); }
int orc_laplacian_ok(void* h) { return static_cast<Laplacian*>(h)->ok ? 1 : 0; }
long orc_laplacian_nnz(void* h) { return static_cast<Laplacian*>(h)->Lp.back(); }
void orc_laplacian_solve(void* h, const double* r, double* d) { laplacian_solve(static_cast<Laplacian*>(h), r, d); }
void orc_laplacian_destroy(void* h) { delete static_cast<Laplacian*>(h); }

void orc_params_1d(int n, const double* g0, const double* dir, double* a, double* b) {
  std::vector<double> gd(2 * size_t(n + 1) * (n + 1));
  vertex_gradients(n, dir, gd.data());
  params_1d(n, g0, gd.data(), a, b);
}

int orc_cubic_roots(double c2, double c1, double c0, double* roots) { return cubic_real_roots(c2, c1, c0, roots); }

double orc_pick_step(const double* z, double* alpha) {
  LineSearch ls = pick_step(z);
  *alpha = ls.alpha;
  return ls.emin;
}

// solve_1D_problem: returns e(alpha)/A; updates xk1, rk1, g0 (in place)
double orc_solve_1d(int n, double* g0, const double* xk, const double* dir, const double* rk, double ek,
                    double* xk1, double* rk1, double* alpha_out, double* z_out) {
  const long F = long(n) * n, V = long(n + 1) * (n + 1);
  const double A = 1.0 / (double(n) * double(n));
  std::vector<double> gd(2 * V), a(F), b(F);
  vertex_gradients(n, dir, gd.data());
  params_1d(n, g0, gd.data(), a.data(), b.data());
  double z[5] = {ek * A, 2. * dot(b.data(), rk, F), dot(b.data(), b.data(), F) + 2 * dot(a.data(), rk, F),
                 2. * dot(a.data(), b.data(), F), dot(a.data(), a.data(), F)};
  LineSearch ls = pick_step(z);
  const double al = ls.alpha;
  for (long f = 0; f < F; ++f) xk1[f] = xk[f] + al * dir[f];
  for (long v = 0; v < 2 * V; ++v) g0[v] = g0[v] + al * gd[v];
  for (long f = 0; f < F; ++f) rk1[f] = a[f] * (al * al) + b[f] * al + rk[f];
  if (alpha_out) *alpha_out = al;
  if (z_out) for (int k = 0; k < 5; ++k) z_out[k] = z[k];
  return ls.emin / A;
}

// compute_conjugate_jacobian_beta  (otlib/otsolver_2dgrid.cpp:248-261)
double orc_cj_beta(int n, const double* rho, const double* xk, const double* rkm1, const double* rk,
                   const double* dhat, const double* dprev, double alpha) {
  (void)dprev;
  const long F = long(n) * n;
  const double eps = alpha / 2.;
  std::vector<double> x(F), reps(F), G;
  for (long f = 0; f < F; ++f) x[f] = xk[f] - eps * dhat[f];
  residual(n, rho, x.data(), reps.data(), G);
  double num = 0, den = 0;
  for (long f = 0; f < F; ++f) {
    const double jd = rk[f] - rkm1[f];
    num += jd * (reps[f] - rk[f]);
    den += jd * jd;
  }
  const double cand = num / den / eps * alpha;
  return (-1. < cand) ? cand : -1.;  // std::max(-1., cand)
}

// GridBasedTransportSolver::solve  (otlib/otsolver_2dgrid.cpp:106-245)
// fwd_xy: SoA forward vertex positions (V x, V y).  Traces hold, per
// iteration, the residual returned by the line search, alpha and beta.
int orc_solve(void* lap, int n, const double* density, int beta_cj, int max_iter, double threshold,
              double max_ratio, double* fwd_xy, double* adj_density, double* res0, double* res_trace,
              double* alpha_trace, double* beta_trace, int trace_cap) {
  const auto* L = static_cast<Laplacian*>(lap);
  if (!L || L->n != n) return -1;
  const long F = long(n) * n, V = long(n + 1) * (n + 1);
  std::vector<double> rho(density, density + F);
  adjust_density(n, rho.data(), max_ratio);
  if (adj_density) std::copy(rho.begin(), rho.end(), adj_density);
  std::vector<double> g0(2 * V, 0.), G;
  std::vector<double> xk(F, 0.), xkp1(F, 0.), rkm1(F, 0.), rk(F, 0.), rkp1(F, 0.), dhat(F, 0.), d(F, 0.);
  double beta = 0, alpha = 0;
  double res = residual(n, rho.data(), xk.data(), rkp1.data(), G);
  if (res0) *res0 = res;
  int it = 0;
  while (it < max_iter && res > threshold) {
    rkm1.swap(rk);
    rk.swap(rkp1);
    laplacian_solve(L, rk.data(), dhat.data());
    if (std::sqrt(dot(dhat.data(), dhat.data(), F)) <= 2 * DBL_MIN) break;
    if (it < 1 || !beta_cj) {
      d = dhat;
    } else {
      beta = orc_cj_beta(n, rho.data(), xk.data(), rkm1.data(), rk.data(), dhat.data(), d.data(), alpha);
      for (long f = 0; f < F; ++f) d[f] = dhat[f] + beta * d[f];
    }
    alpha = 0;
    res = orc_solve_1d(n, g0.data(), xk.data(), d.data(), rk.data(), res, xkp1.data(), rkp1.data(), &alpha, nullptr);
    xk.swap(xkp1);
    if (it < trace_cap) {
      if (res_trace) res_trace[it] = res;
      if (alpha_trace) alpha_trace[it] = alpha;
      if (beta_trace) beta_trace[it] = beta;
    }
    ++it;
  }
  std::vector<double> Gf(2 * V);
  vertex_gradients(n, xk.data(), Gf.data());
  for (long v = 0; v < V; ++v) {
    const P2 o = origin_point(n, int(v));
    fwd_xy[v] = o.x + Gf[v];
    fwd_xy[V + v] = o.y + Gf[V + v];
  }
  return it;
}

void* orc_bvh_build(int n, const double* xy, int leaf_size, int max_depth) {
  return bvh_build(n, xy, leaf_size, max_depth);
}
void orc_bvh_destroy(void* h) { delete static_cast<BVH*>(h); }
int orc_bvh_num_nodes(void* h) { return int(static_cast<BVH*>(h)->nodes.size()); }

// Flattened BVH export: nodes as (lo.x, lo.y, hi.x, hi.y) + (first, count, leaf), face order.
void orc_bvh_export(void* h, double* boxes, int* meta, int* faces) {
  const BVH* B = static_cast<BVH*>(h);
  for (size_t k = 0; k < B->nodes.size(); ++k) {
    const BNode& nd = B->nodes[k];
    boxes[4 * k + 0] = nd.lo[0]; boxes[4 * k + 1] = nd.lo[1];
    boxes[4 * k + 2] = nd.hi[0]; boxes[4 * k + 3] = nd.hi[1];
    meta[3 * k + 0] = nd.first; meta[3 * k + 1] = nd.count; meta[3 * k + 2] = nd.leaf;
  }
  std::copy(B->faces.begin(), B->faces.end(), faces);
}

// apply_inverse_map over m interleaved points; out interleaved; face = chosen
// face (-1 when none), nhits = number of overlapping faces found.
void orc_inverse(void* h, const double* density, const double* pts, long m, int fast, double* out, int* face,
                 int* nhits) {
  const BVH* B = static_cast<BVH*>(h);
  std::vector<Hit> hits;
  for (long k = 0; k < m; ++k) {
    int f;
    const int nh = inverse_point(*B, density, {pts[2 * k], pts[2 * k + 1]}, fast != 0, out + 2 * k, &f, hits);
    if (face) face[k] = f;
    if (nhits) nhits[k] = nh;
  }
}

// brute force: every face containing q under the reference predicates, in face order
int orc_locate_all(int n, const double* xy, double qx, double qy, int* faces, int cap) {
  const long V = long(n + 1) * (n + 1);
  int k = 0;
  for (int f = 0; f < n * n; ++f) {
    int v[4];
    quad_vertices(n, f, v);
    P2 p[4];
    for (int t = 0; t < 4; ++t) p[t] = {xy[v[t]], xy[V + v[t]]};
    if (!inside_quad({qx, qy}, p)) continue;
    double w[4];
    if (!bilinear_weights({qx, qy}, p, w)) continue;
    bool ok = true;
    for (int t = 0; t < 4; ++t) ok = ok && w[t] <= 1.0f && w[t] >= 0.0f;
    if (ok && k < cap) faces[k++] = f;
  }
  return k;
}

// rasterize_image, per-face density mode  (otlib/utils/rasterizer.cpp:138-237)
// img row-major rows x cols; density may be null (all ones).
// Where the bilinear solve fails (D < 0) the reference ignores the return
// value and tests an uninitialised Eigen::Vector4d (rasterizer.cpp:169-171).
// In the reference's Release build that vector occupies the same stack slot
// on every pixel and every face, so it holds the weights of the last pixel
// whose solve succeeded, in scan order (faces in index order, rows y, then
// columns x).  That observed behaviour is reproduced here (it matches the
// reference's images bit for bit); before the first successful solve the
// weights count as "not covered".
void orc_rasterize(int n, const double* xy, const double* density, int rows, int cols, double* img) {
  const long V = long(n + 1) * (n + 1), F = long(n) * n;
  std::fill(img, img + long(rows) * cols, 0.);
  const float scale = float(1. / double(F));
  double w[4] = {-1., -1., -1., -1.};  // carried across pixels and faces (see above)
  for (long f = 0; f < F; ++f) {
    int v[4];
    quad_vertices(n, int(f), v);
    P2 p[4];
    for (int t = 0; t < 4; ++t) p[t] = {xy[v[t]], xy[V + v[t]]};
    const double mass = double(scale) * (density ? density[f] : 1.0);
    const double area = std::abs(sarea(p[0], p[1], p[2]) + sarea(p[0], p[2], p[3]));
    double lo[2] = {p[0].x, p[0].y}, hi[2] = {p[0].x, p[0].y};
    for (int t = 1; t < 4; ++t) {
      lo[0] = std::min(lo[0], p[t].x); lo[1] = std::min(lo[1], p[t].y);
      hi[0] = std::max(hi[0], p[t].x); hi[1] = std::max(hi[1], p[t].y);
    }
    const int x0 = std::max(int(lo[0] * double(cols)), 0), y0 = std::max(int(lo[1] * double(rows)), 0);
    const int x1 = std::min(int(hi[0] * double(cols)) + 1, cols), y1 = std::min(int(hi[1] * double(rows)) + 1, rows);
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        const P2 c = {(double(x) + 0.5) / double(cols), (double(y) + 0.5) / double(rows)};
        bilinear_weights(c, p, w);  // leaves w unchanged when D < 0
        bool ok = true;
        for (int t = 0; t < 4; ++t) ok = ok && w[t] <= 1.0f && w[t] >= 0.0f;
        if (ok) img[long(y) * cols + x] = mass / area;
      }
  }
}

// How often the carried weights of orc_rasterize decide a pixel, on a given
// mesh (same scan as orc_rasterize).  out[0] pixels scanned, out[1] pixels
// whose bilinear solve fails (D < 0), out[2] of those, weights carried from
// an earlier pixel of the same face, out[3] carried from a preceding face,
// out[4] no successful solve before them (not covered), out[5] failing pixels
// that end up covered (carried weights all in [0, 1]), out[6] the largest
// number of faces between a failing pixel's face and the face whose solve
// its weights come from (the device looks back at most 64 faces).
void orc_raster_stats(int n, const double* xy, int rows, int cols, long long* out) {
  const long V = long(n + 1) * (n + 1), F = long(n) * n;
  for (int k = 0; k < 7; ++k) out[k] = 0;
  double w[4] = {-1., -1., -1., -1.};
  long last_ok_face = -1;
  for (long f = 0; f < F; ++f) {
    int v[4];
    quad_vertices(n, int(f), v);
    P2 p[4];
    for (int t = 0; t < 4; ++t) p[t] = {xy[v[t]], xy[V + v[t]]};
    double lo[2] = {p[0].x, p[0].y}, hi[2] = {p[0].x, p[0].y};
    for (int t = 1; t < 4; ++t) {
      lo[0] = std::min(lo[0], p[t].x); lo[1] = std::min(lo[1], p[t].y);
      hi[0] = std::max(hi[0], p[t].x); hi[1] = std::max(hi[1], p[t].y);
    }
    const int x0 = std::max(int(lo[0] * double(cols)), 0), y0 = std::max(int(lo[1] * double(rows)), 0);
    const int x1 = std::min(int(hi[0] * double(cols)) + 1, cols), y1 = std::min(int(hi[1] * double(rows)) + 1, rows);
    for (int y = y0; y < y1; ++y)
      for (int x = x0; x < x1; ++x) {
        ++out[0];
        const P2 c = {(double(x) + 0.5) / double(cols), (double(y) + 0.5) / double(rows)};
        if (bilinear_weights(c, p, w)) {
          last_ok_face = f;
          continue;
        }
        ++out[1];
        if (last_ok_face < 0) {
          ++out[4];
          continue;
        }
        if (last_ok_face == f) {
          ++out[2];
        } else {
          ++out[3];
          out[6] = std::max<long long>(out[6], f - last_ok_face);
        }
        bool ok = true;
        for (int t = 0; t < 4; ++t) ok = ok && w[t] <= 1.0f && w[t] >= 0.0f;
        if (ok) ++out[5];
      }
  }
}

// transport_cost on a quad grid mesh  (otlib/transport_map.cpp:131-215).
// With one density value per quad the reference takes its triangle branch on
// every quad: the triangle (v3, v0, v1) given by the face's first halfedge.
double orc_transport_cost(int n, const double* src, const double* dst, const double* density, double* cost) {
  const long V = long(n + 1) * (n + 1), F = long(n) * n;
  double res = 0, integral = 0;
  for (long f = 0; f < F; ++f) {
    int v[4];
    quad_vertices(n, int(f), v);
    const int a = v[3], b = v[0], c = v[1];
    const P2 sa = {src[a], src[V + a]}, sb = {src[b], src[V + b]}, sc = {src[c], src[V + c]};
    const double area = sarea(sa, sb, sc);
    const P2 da = {dst[a] - sa.x, dst[V + a] - sa.y};
    const P2 db = {dst[b] - sb.x, dst[V + b] - sb.y};
    const P2 dc = {dst[c] - sc.x, dst[V + c] - sc.y};
    auto d2 = [&](double l1, double l2, double l3) {
      const double x = l1 * da.x + l2 * db.x + l3 * dc.x;
      const double y = l1 * da.y + l2 * db.y + l3 * dc.y;
      return x * x + y * y;
    };
    double fc = (d2(0.5, 0.5, 0) + d2(0, 0.5, 0.5) + d2(0.5, 0, 0.5)) / 3.0;
    integral += density[f] * area;
    fc = density[f] * area * fc;
    res += fc;
    if (cost) cost[f] = fc;
  }
  if (cost) for (long f = 0; f < F; ++f) cost[f] /= integral;
  return res / integral;
}

// compute_transport_cost  (otlib/otsolver_2dgrid.cpp:608-640); G = vertex displacements (SoA)
double orc_solver_transport_cost(int n, const double* G, const double* rho, double* cost) {
  const long V = long(n + 1) * (n + 1);
  const double A = 1.0 / (double(n) * double(n));
  const double z1 = std::sqrt(1. / 3.) / 2. + 0.5, z2 = -std::sqrt(1. / 3.) / 2. + 0.5;
  double total = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const int v1 = vidx(n, i, j), v2 = vidx(n, i + 1, j), v3 = vidx(n, i + 1, j + 1), v4 = vidx(n, i, j + 1);
      auto d2 = [&](double u, double v) {
        const double x = (1 - u) * (1 - v) * G[v1] + u * (1 - v) * G[v2] + u * v * G[v3] + (1 - u) * v * G[v4];
        const double y = (1 - u) * (1 - v) * G[V + v1] + u * (1 - v) * G[V + v2] + u * v * G[V + v3] +
                         (1 - u) * v * G[V + v4];
        return x * x + y * y;
      };
      const double c = rho[fidx(n, i, j)] * A * (d2(z1, z1) + d2(z1, z2) + d2(z2, z2) + d2(z2, z1)) / 4.;
      if (cost) cost[fidx(n, i, j)] = c;
      total += c;
    }
  return total;
}

}  // extern "C"
