// Container-side driver around the compiled reference (ggael/otmap, built from its
// own sources under /root/reference by oracle/Makefile into oracle/_ref/).
//
// This file is test infrastructure: it exposes the reference's own solver,
// transport-map, BVH and rasteriser entry points through a flat C ABI so that
// tests/golden/gen_golden.py can record golden vectors and bench.py can time
// the reference's CPU path.  Nothing in the product links it.
//
// Reference entry points wrapped (paths relative to the reference root):
//   GridBasedTransportSolver::init/solve       otlib/otsolver_2dgrid.cpp:75-245
//   compute_vertex_gradients / compute_residual otlib/otsolver_2dgrid.cpp:362-465
//   compute_1D_problem_parameters / solve_1D    otlib/otsolver_2dgrid.cpp:467-606
//   compute_conjugate_jacobian_beta             otlib/otsolver_2dgrid.cpp:248-261
//   TransportMap::inv / inv_fast, BVH2D         otlib/transport_map.cpp:50-102, otlib/utils/bvh2d.cpp
//   rasterize_image                             otlib/utils/rasterizer.cpp:188-308
//   transport_cost                              otlib/transport_map.cpp:131-215
//   compute_transport_cost                      otlib/otsolver_2dgrid.cpp:608-640
//   nestdiss_ordering                           otlib/details/nested_dissection.cpp:54-61
//   load_point_cloud_dat, generate_blue_noise_tile apps/common/image_utils.cpp:148-173,209-231

#include "otsolver_2dgrid.h"
#include "transport_map.h"
#include "utils/bvh2d.h"
#include "utils/mesh_utils.h"
#include "utils/rasterizer.h"
#include "details/nested_dissection.h"
#include "common/image_utils.h"

#include <xmmintrin.h>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <time.h>

using namespace Eigen;
using surface_mesh::Surface_mesh;

namespace {

// The reference constructor sets FTZ/DAZ on the calling thread; keep the
// host process's floating-point mode untouched across every call.
struct MxcsrGuard {
  unsigned int saved;
  MxcsrGuard() : saved(_mm_getcsr()) {}
  ~MxcsrGuard() { _mm_setcsr(saved); }
};

// Gives access to the protected stencils and the factorised Laplacian.
struct Probe : public otmap::GridBasedTransportSolver {
  Probe() { m_gridSize = 0; }
  void set_density(const VectorXd* d) { m_input_density = d; }
  VectorXd lap_solve(const VectorXd& r) {
    VectorXd d = m_laplacian_solver.solve(r);
    d.array() -= d.mean();
    return d;
  }
  void grads(const VectorXd& psi, MatrixX2d& g) const { compute_vertex_gradients(psi, g); }
  double residual(const VectorXd& psi, VectorXd& out) const {
    return compute_residual(psi, out);
  }
  void params(const MatrixX2d& g0, const VectorXd& psi, const VectorXd& dir, VectorXd& a, VectorXd& b) {
    m_cache_1D_g0 = g0;
    compute_1D_problem_parameters(psi, dir, a, b);
  }
  double solve1d(const MatrixX2d& g0, const VectorXd& xk, const VectorXd& dir, const VectorXd& rk, double ek,
                 VectorXd& xk1, VectorXd& rk1, MatrixX2d& g0_out, double* alpha) {
    m_cache_1D_g0 = g0;
    double e = solve_1D_problem(xk, dir, rk, ek, xk1, rk1, alpha);
    g0_out = m_cache_1D_g0;
    return e;
  }
  double beta(const VectorXd& xk, const VectorXd& rkm1, const VectorXd& rk, const VectorXd& dh,
              const VectorXd& dp, double alpha) const {
    return compute_conjugate_jacobian_beta(xk, rkm1, rk, dh, dp, alpha);
  }
  void adjust(VectorXd& d, double max_ratio) { adjust_density(d, max_ratio); }
  void transport_cost(const MatrixX2d& g, VectorXd& cost) const { compute_transport_cost(g, cost); }
};

std::map<int, std::unique_ptr<Probe>>& probes() {
  static std::map<int, std::unique_ptr<Probe>> p;
  return p;
}

Probe& probe(int n) {
  auto& m = probes();
  auto it = m.find(n);
  if (it != m.end()) return *it->second;
  std::unique_ptr<Probe> p(new Probe);
  p->set_verbose_level(0);
  p->init(n);
  Probe& ref = *p;
  m[n] = std::move(p);
  return ref;
}

double now_s() {
  timespec ts;
  clock_gettime(CLOCK_REALTIME, &ts);
  return double(ts.tv_sec) + 1e-9 * double(ts.tv_nsec);
}

std::shared_ptr<Surface_mesh> grid_mesh(int n, const double* xy) {
  auto m = std::make_shared<Surface_mesh>();
  otmap::generate_quad_mesh(n + 1, n + 1, *m);
  if (xy) {
    const int nv = (n + 1) * (n + 1);
    for (int v = 0; v < nv; ++v) m->points()[v] = Vector2d(xy[v], xy[nv + v]);
  }
  return m;
}

// gives access to the lazily built BVH of a TransportMap
struct MapProbe : public otmap::TransportMap {
  using otmap::TransportMap::TransportMap;
  const otmap::BVH2D* bvh() const { return m_bvh; }
};

}  // namespace

extern "C" {

// Full solve through the reference's public API.  Traces (residual after each
// iteration, alpha, beta) are parsed from the reference's own verbose-5 log,
// printed at 17 significant digits.  fwd_xy is SoA: x of all V vertices, then y.
int ref_solve(int n, const double* density, int beta_cj, int max_iter, double threshold, double max_ratio,
              double* fwd_xy, double* adj_density, int* iters, double* res0,
              double* res_trace, double* alpha_trace, double* beta_trace, int trace_cap,
              int traced, double* t_init, double* t_solve) {
  MxcsrGuard guard;
  const int F = n * n, V = (n + 1) * (n + 1);
  otmap::GridBasedTransportSolver solver;
  solver.set_verbose_level(traced ? 5 : 0);
  std::ostringstream log;
  std::streambuf* old = nullptr;
  if (traced) {
    old = std::cout.rdbuf(log.rdbuf());
    std::cout << std::setprecision(17);
  }
  double t0 = now_s();
  // m_gridSize is read before it is set in init(); a fresh solver is used per call
  solver.init(n);
  double t1 = now_s();
  otmap::SolverOptions opt;
  opt.beta = beta_cj ? otmap::BetaOpt::ConjugateJacobian : otmap::BetaOpt::Zero;
  opt.max_iter = max_iter;
  opt.threshold = threshold;
  opt.max_ratio = max_ratio;
  Map<const VectorXd> dens(density, F);
  otmap::TransportMap tm = solver.solve(dens, opt);
  double t2 = now_s();
  if (traced) {
    std::cout.rdbuf(old);
    std::cout << std::setprecision(6);
  }
  if (t_init) *t_init = t1 - t0;
  if (t_solve) *t_solve = t2 - t1;
  const auto& pts = tm.fwd_mesh().get_vertex_property<surface_mesh::Point>("v:point").vector();
  for (int v = 0; v < V; ++v) { fwd_xy[v] = pts[v].x(); fwd_xy[V + v] = pts[v].y(); }
  if (adj_density) for (int f = 0; f < F; ++f) adj_density[f] = tm.density()(f);
  int count = 0;
  if (traced) {
    std::istringstream in(log.str());
    std::string line;
    while (std::getline(in, line)) {
      size_t p0 = line.find("initial L2=");
      if (p0 != std::string::npos && res0) *res0 = std::stod(line.substr(p0 + 11));
      size_t pa = line.find("alpha=");
      if (pa == std::string::npos) continue;
      size_t pb = line.find("beta=");
      size_t pr = line.find("res^2=");
      if (count < trace_cap) {
        alpha_trace[count] = std::stod(line.substr(pa + 6));
        beta_trace[count] = std::stod(line.substr(pb + 5));
        res_trace[count] = std::stod(line.substr(pr + 6));
      }
      ++count;
    }
    if (iters) *iters = count;
  } else if (iters) {
    *iters = -1;
  }
  return 0;
}

// Timing of the reference's solve(): one solver, init(n) once, then `reps`
// solves of the same density (verbose 0); t_solve[k] = wall seconds of solve k.
int ref_solve_timed(int n, const double* density, int reps, double* t_init, double* t_solve) {
  MxcsrGuard guard;
  otmap::GridBasedTransportSolver solver;
  solver.set_verbose_level(0);
  double t0 = now_s();
  solver.init(n);
  if (t_init) *t_init = now_s() - t0;
  Map<const VectorXd> dens(density, n * n);
  for (int k = 0; k < reps; ++k) {
    t0 = now_s();
    otmap::TransportMap tm = solver.solve(dens, otmap::SolverOptions());
    t_solve[k] = now_s() - t0;
  }
  return 0;
}

void ref_adjust_density(int n, double* rho, double max_ratio) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  VectorXd d = Map<VectorXd>(rho, n * n);
  p.adjust(d, max_ratio);
  Map<VectorXd>(rho, n * n) = d;
}

void ref_vertex_gradients(int n, const double* psi, double* G) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  MatrixX2d g;
  p.grads(Map<const VectorXd>(psi, n * n), g);
  const int V = (n + 1) * (n + 1);
  for (int v = 0; v < V; ++v) { G[v] = g(v, 0); G[V + v] = g(v, 1); }
}

double ref_residual(int n, const double* rho_adj, const double* psi, double* r) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  VectorXd d = Map<const VectorXd>(rho_adj, n * n);
  p.set_density(&d);
  VectorXd out(n * n);
  double e = p.residual(Map<const VectorXd>(psi, n * n), out);
  Map<VectorXd>(r, n * n) = out;
  p.set_density(nullptr);
  return e;
}

void ref_laplacian_solve(int n, const double* r, double* dhat) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  VectorXd d = p.lap_solve(Map<const VectorXd>(r, n * n));
  Map<VectorXd>(dhat, n * n) = d;
}

static MatrixX2d soa_to_grads(int n, const double* G) {
  const int V = (n + 1) * (n + 1);
  MatrixX2d g(V, 2);
  for (int v = 0; v < V; ++v) { g(v, 0) = G[v]; g(v, 1) = G[V + v]; }
  return g;
}

void ref_1d_params(int n, const double* g0, const double* dir, double* a, double* b) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  VectorXd va(n * n), vb(n * n);
  VectorXd psi = VectorXd::Zero(n * n);
  p.params(soa_to_grads(n, g0), psi, Map<const VectorXd>(dir, n * n), va, vb);
  Map<VectorXd>(a, n * n) = va;
  Map<VectorXd>(b, n * n) = vb;
}

double ref_solve_1d(int n, const double* g0, const double* xk, const double* dir, const double* rk, double ek,
                    double* xk1, double* rk1, double* g0_out, double* alpha) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  const int F = n * n, V = (n + 1) * (n + 1);
  VectorXd x1(F), r1(F);
  MatrixX2d go;
  double e = p.solve1d(soa_to_grads(n, g0), Map<const VectorXd>(xk, F), Map<const VectorXd>(dir, F),
                       Map<const VectorXd>(rk, F), ek, x1, r1, go, alpha);
  Map<VectorXd>(xk1, F) = x1;
  Map<VectorXd>(rk1, F) = r1;
  for (int v = 0; v < V; ++v) { g0_out[v] = go(v, 0); g0_out[V + v] = go(v, 1); }
  return e;
}

double ref_cj_beta(int n, const double* rho_adj, const double* xk, const double* rkm1, const double* rk,
                   const double* dhat, const double* dprev, double alpha) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  const int F = n * n;
  VectorXd d = Map<const VectorXd>(rho_adj, F);
  p.set_density(&d);
  double b = p.beta(Map<const VectorXd>(xk, F), Map<const VectorXd>(rkm1, F), Map<const VectorXd>(rk, F),
                    Map<const VectorXd>(dhat, F), Map<const VectorXd>(dprev, F), alpha);
  p.set_density(nullptr);
  return b;
}

void ref_nestdiss(int n, int* perm) { otmap::nestdiss_ordering(n, perm); }

// Inverse map of m points (interleaved xy).  For every point: the result of
// TransportMap::inv (fast==0) or inv_fast (fast==1), the number of BVH hits and
// the face id of the first hit in traversal order.  Points with no hit (where
// the reference reads an empty vector) get NaN and nhits = 0.
int ref_inverse(int n, const double* fwd_xy, const double* density, const double* pts, long m, int fast,
                double* out, int* first_face, int* nhits, double* t_build, double* t_query) {
  MxcsrGuard guard;
  auto origin = grid_mesh(n, nullptr);
  auto fwd = grid_mesh(n, fwd_xy);
  auto dens = std::make_shared<VectorXd>(Map<const VectorXd>(density, n * n));
  MapProbe tm(origin, fwd, dens);
  double t0 = now_s();
  tm.init_inverse();
  double t1 = now_s();
  std::vector<otmap::BVH2D::Hit> hits;
  for (long k = 0; k < m; ++k) {
    Vector2d p(pts[2 * k], pts[2 * k + 1]);
    Vector2d pc = p.array().max(0.).min(1.);
    hits.clear();
    tm.bvh()->query_all(pc, hits);
    if (nhits) nhits[k] = int(hits.size());
    if (first_face) first_face[k] = hits.empty() ? -1 : hits[0].face_id.idx();
    if (hits.empty()) {
      out[2 * k] = out[2 * k + 1] = std::numeric_limits<double>::quiet_NaN();
      continue;
    }
    Vector2d r = fast ? tm.inv_fast(p) : tm.inv(p);
    out[2 * k] = r.x();
    out[2 * k + 1] = r.y();
  }
  double t2 = now_s();
  if (t_build) *t_build = t1 - t0;
  if (t_query) *t_query = t2 - t1;
  return 0;
}

// Timing-only batch inverse through the reference's apply_inverse_map
// (lazy BVH build + serial inv() loop).  pts is updated in place.
int ref_apply_inverse_map(int n, const double* fwd_xy, const double* density, double* pts, long m, double* t_total) {
  MxcsrGuard guard;
  auto origin = grid_mesh(n, nullptr);
  auto fwd = grid_mesh(n, fwd_xy);
  auto dens = std::make_shared<VectorXd>(Map<const VectorXd>(density, n * n));
  otmap::TransportMap tm(origin, fwd, dens);
  std::vector<Vector2d> p(m);
  for (long k = 0; k < m; ++k) p[k] = Vector2d(pts[2 * k], pts[2 * k + 1]);
  double t0 = now_s();
  otmap::apply_inverse_map(tm, p, 0);
  double t1 = now_s();
  for (long k = 0; k < m; ++k) { pts[2 * k] = p[k].x(); pts[2 * k + 1] = p[k].y(); }
  if (t_total) *t_total = t1 - t0;
  return 0;
}

// Rasterise the (n+1)^2-vertex quad mesh into a rows x cols image; img is
// row-major (img[y*cols+x] == reference img(y,x)).  density may be null (ones).
void ref_rasterize(int n, const double* mesh_xy, const double* density, int rows, int cols, double* img) {
  MxcsrGuard guard;
  auto mesh = grid_mesh(n, mesh_xy);
  MatrixXd im(rows, cols);
  if (density)
    otmap::rasterize_image(*mesh, Map<const VectorXd>(density, n * n), im);
  else
    otmap::rasterize_image(*mesh, im);
  for (int y = 0; y < rows; ++y)
    for (int x = 0; x < cols; ++x) img[y * cols + x] = im(y, x);
}

double ref_transport_cost(int n, const double* src_xy, const double* dst_xy, const double* density,
                          double* cost_per_face) {
  MxcsrGuard guard;
  auto src = grid_mesh(n, src_xy);
  auto dst = grid_mesh(n, dst_xy);
  VectorXd c;
  double r = otmap::transport_cost(*src, *dst, Map<const VectorXd>(density, n * n), cost_per_face ? &c : nullptr);
  if (cost_per_face) Map<VectorXd>(cost_per_face, n * n) = c;
  return r;
}

// GridBasedTransportSolver::compute_transport_cost on vertex displacements
// disp_xy (SoA, the gradients the solver passes) and the adjusted density the
// solver would hold; returns the sum of the per-face costs.
double ref_solver_transport_cost(int n, const double* disp_xy, const double* rho_adj, double* cost_per_face) {
  MxcsrGuard guard;
  Probe& p = probe(n);
  VectorXd d = Map<const VectorXd>(rho_adj, n * n);
  p.set_density(&d);
  VectorXd c;
  p.transport_cost(soa_to_grads(n, disp_xy), c);
  p.set_density(nullptr);
  if (cost_per_face) Map<VectorXd>(cost_per_face, n * n) = c;
  return c.sum();
}

// The reference apps' blue-noise tile reader and tiler: the tile as read
// from `path` (interleaved, up to cap points; returns the count), and the
// points of generate_blue_noise_tile(n) from that tile (up to cap; returns
// the count).
long ref_load_point_cloud_dat(const char* path, double* out, long cap) {
  std::vector<Vector2d> pts;
  if (!load_point_cloud_dat(path, pts)) return -1;
  for (long k = 0; k < long(pts.size()) && k < cap; ++k) { out[2 * k] = pts[k].x(); out[2 * k + 1] = pts[k].y(); }
  return long(pts.size());
}

long ref_blue_noise_tile(int n, const double* tile, long nb, double* out, long cap) {
  std::vector<Vector2d> t(nb), pts;
  for (long k = 0; k < nb; ++k) t[k] = Vector2d(tile[2 * k], tile[2 * k + 1]);
  generate_blue_noise_tile(n, pts, t);
  for (long k = 0; k < long(pts.size()) && k < cap; ++k) { out[2 * k] = pts[k].x(); out[2 * k + 1] = pts[k].y(); }
  return long(pts.size());
}

}  // extern "C"
