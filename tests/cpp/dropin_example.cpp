// A caller of the reference's C++ API (otlib/otsolver_2dgrid.h,
// otlib/transport_map.h, otlib/utils/rasterizer.h), compiled against the
// drop-in headers include/otmap/*.h without Eigen and linked with
// libotmap_amd.so.  Mirrors the solve -> inverse -> rasterize sequence of the
// reference's apps/otmap.cpp and apps/stippling.cpp on raw binary files so
// that a test can compare every output with the oracle:
//
//   dropin_example <n> <density.f64> <points.f64> <out_prefix> [mixed]
//
// writes <out_prefix>.fwd.f64 (2*(n+1)^2 SoA positions), .dens.f64 (n*n),
// .inv.f64 (points replaced by apply_inverse_map), .img.f64 (n x n raster,
// row-major), .stats.txt (iterations and residuals).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "otmap/otsolver_2dgrid.h"
#include "otmap/rasterizer.h"
#include "otmap/transport_map.h"

using namespace otmap;

static std::vector<double> read_f64(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  std::vector<double> v(size_t(f.tellg()) / sizeof(double));
  f.seekg(0);
  f.read(reinterpret_cast<char*>(v.data()), std::streamsize(v.size() * sizeof(double)));
  return v;
}

static bool write_f64(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(double)));
  return bool(f);
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cerr << "usage: " << argv[0] << " n density.f64 points.f64 out_prefix [mixed]\n";
    return 2;
  }
  const int n = std::stoi(argv[1]);
  const std::vector<double> density = read_f64(argv[2]);
  const std::vector<double> raw_pts = read_f64(argv[3]);
  const std::string out = argv[4];
  const bool mixed = argc > 5 && std::string(argv[5]) == "mixed";

  GridBasedTransportSolver solver;
  solver.set_verbose_level(0);
  solver.init(n);

  SolverOptions opt;
  opt.beta = BetaOpt::ConjugateJacobian;
  opt.max_iter = 1000;
  opt.threshold = 1e-7;
  if (mixed) {
    opt.precision = Precision::Mixed;
    opt.mixed_min_n = 0;  // the BF16x3 transform at this size too
  }
  TransportMap tmap = solver.solve(density, opt);
  if (!tmap.valid()) return 1;

  std::vector<Point> pts(raw_pts.size() / 2);
  for (size_t k = 0; k < pts.size(); ++k) pts[k] = Point(raw_pts[2 * k], raw_pts[2 * k + 1]);
  apply_inverse_map(tmap, pts, 0);
  std::vector<double> inv(2 * pts.size());
  for (size_t k = 0; k < pts.size(); ++k) {
    inv[2 * k] = pts[k].x();
    inv[2 * k + 1] = pts[k].y();
  }

  Image img(n, n);
  rasterize_image(tmap.fwd_mesh(), tmap.density(), img);
  std::vector<double> img_rm(size_t(n) * n);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) img_rm[size_t(y) * n + x] = img(y, x);

  const otm_stats& st = solver.last_stats();
  std::ofstream stats(out + ".stats.txt");
  stats.precision(17);
  stats << st.iterations << " " << st.initial_residual << " " << st.final_residual << "\n";

  const bool ok = write_f64(out + ".fwd.f64", tmap.fwd_mesh().xy()) && write_f64(out + ".dens.f64", tmap.density()) &&
                  write_f64(out + ".inv.f64", inv) && write_f64(out + ".img.f64", img_rm) &&
                  tmap.fwd_mesh().write(out + ".fwd.off");
  std::printf("dropin_example: n=%d iterations=%d residual=%.3e\n", n, st.iterations, st.final_residual);
  return ok ? 0 : 1;
}
