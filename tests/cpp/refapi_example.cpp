// A caller written against the reference's own signatures (ggael/otmap
// otlib/otsolver_2dgrid.h, otlib/transport_map.h, otlib/utils/rasterizer.h,
// surface_mesh::Surface_mesh, Eigen types), compiled unchanged against the
// drop-in headers: the include path points at include/otmap (where the
// reference's points at otlib) and Eigen is on it, as in the reference's own
// build.  The body follows apps/barycenters.cpp:167-290 for two densities and
// the cost report of apps/otmap.cpp: solve each density (solve(vec(img))),
// inverse meshes through apply_inverse_map on the origin mesh's points(),
// the halfway interpolation by Vertex handles, rasterize_image into a
// MatrixXd scaled to the interpolated mean, and transport_cost.
//
//   refapi_example <n> <density0.f64> <density1.f64> <out_prefix>
//
// writes <out_prefix>.fwd{0,1}.f64 and .inv{0,1}.f64 (SoA positions),
// .mid.f64 (the interpolated mesh), .img.f64 (n x n, row-major), .cost.f64
// (total, then the per-face costs of map 0).
#include <Eigen/Core>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "otsolver_2dgrid.h"
#include "surface_mesh/Surface_mesh.h"
#include "transport_map.h"
#include "utils/rasterizer.h"

using namespace Eigen;
using namespace otmap;
using surface_mesh::Surface_mesh;

// apps/common/eigen_addons.h: an image as the solver's face-order vector
inline Map<const VectorXd> vec(const MatrixXd& m) { return Map<const VectorXd>(m.data(), m.size()); }

template <typename T, typename S>
T lerp(S u, const T& a0, const T& a1) {
  return (1. - u) * a0 + u * a1;
}

static MatrixXd read_image(const std::string& path, int n) {
  MatrixXd m(n, n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(m.data()), std::streamsize(sizeof(double) * n * n));
  return m;
}

static void write_xy(const std::string& path, const Surface_mesh& mesh) {
  std::vector<double> xy(2 * mesh.vertices_size());
  const size_t V = mesh.vertices_size();
  for (size_t v = 0; v < V; ++v) {
    xy[v] = mesh.points()[v].x();
    xy[V + v] = mesh.points()[v].y();
  }
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(xy.data()), std::streamsize(xy.size() * sizeof(double)));
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cerr << "usage: " << argv[0] << " n density0.f64 density1.f64 out_prefix\n";
    return 2;
  }
  const int n = std::stoi(argv[1]);
  const std::string out = argv[4];
  std::vector<MatrixXd> inputs = {read_image(argv[2], n), read_image(argv[3], n)};

  GridBasedTransportSolver otsolver;
  otsolver.set_verbose_level(0);
  SolverOptions opt;
  std::vector<TransportMap> tmaps;
  std::vector<double> density_means;
  for (const MatrixXd& img : inputs) {
    otsolver.init(img.rows());
    tmaps.push_back(otsolver.solve(vec(img), opt));
    density_means.push_back(img.mean());
  }

  std::vector<Surface_mesh> inv_maps(tmaps.size());
  for (size_t k = 0; k < tmaps.size(); ++k) {
    inv_maps[k] = tmaps[k].origin_mesh();
    apply_inverse_map(tmaps[k], inv_maps[k].points(), 0);
    write_xy(out + ".fwd" + std::to_string(k) + ".f64", tmaps[k].fwd_mesh());
    write_xy(out + ".inv" + std::to_string(k) + ".f64", inv_maps[k]);
  }

  const double alpha = 0.5;
  Surface_mesh result = inv_maps[0];
  const int nv = result.vertices_size();
  for (int j = 0; j < nv; ++j) {
    Surface_mesh::Vertex v(j);
    result.position(v) = lerp(alpha, inv_maps[0].position(v), inv_maps[1].position(v));
  }
  write_xy(out + ".mid.f64", result);

  MatrixXd img(n, n);
  rasterize_image(result, img);
  img = img * (lerp(alpha, density_means[0], density_means[1]) / img.mean());
  std::vector<double> rm(size_t(n) * n);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) rm[size_t(y) * n + x] = img(y, x);
  std::ofstream(out + ".img.f64", std::ios::binary)
      .write(reinterpret_cast<const char*>(rm.data()), std::streamsize(rm.size() * sizeof(double)));

  VectorXd cost_per_face;
  const double cost = transport_cost(tmaps[0].origin_mesh(), tmaps[0].fwd_mesh(), tmaps[0].density(), &cost_per_face);
  std::vector<double> c(1 + cost_per_face.size());
  c[0] = cost;
  for (long f = 0; f < cost_per_face.size(); ++f) c[1 + f] = cost_per_face(f);
  std::ofstream(out + ".cost.f64", std::ios::binary)
      .write(reinterpret_cast<const char*>(c.data()), std::streamsize(c.size() * sizeof(double)));
  std::cout << "refapi_example: n=" << n << " cost=" << cost << "\n";
  return 0;
}
