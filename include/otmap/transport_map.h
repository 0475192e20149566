// Drop-in for the reference's otlib/transport_map.h and the subset of
// surface_mesh::Surface_mesh its callers use on a regular quad grid
// (ggael/otmap).  Header-only over the C ABI of libotmap_amd.so.
//
// GridMesh: (n+1)^2 vertices, vertex (i, j) -> index j + i*(n+1); n^2 quads,
// quad (i, j) -> index j + i*n with vertices (i,j), (i+1,j), (i+1,j+1),
// (i,j+1) (otlib/utils/mesh_utils.cpp:90-122).  Positions are stored SoA.
#pragma once

#include <array>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../otmap_capi.h"

namespace otmap {

struct Point2 {
  double x = 0., y = 0.;
};

class GridMesh {
 public:
  GridMesh() = default;
  GridMesh(int n, std::vector<double> xy) : n_(n), xy_(std::move(xy)) {}

  // generate_quad_mesh(n+1, n+1): vertex (i, j) at (i*(1/n), j*(1/n))
  static GridMesh regular(int n) {
    const size_t V = size_t(n + 1) * (n + 1);
    std::vector<double> xy(2 * V);
    const double dx = 1. / double(n);
    for (int i = 0; i <= n; ++i)
      for (int j = 0; j <= n; ++j) {
        xy[j + size_t(i) * (n + 1)] = double(i) * dx;
        xy[V + j + size_t(i) * (n + 1)] = double(j) * dx;
      }
    return GridMesh(n, std::move(xy));
  }

  int grid_size() const { return n_; }
  unsigned vertices_size() const { return unsigned((n_ + 1) * (n_ + 1)); }
  unsigned faces_size() const { return unsigned(n_ * n_); }
  unsigned n_vertices() const { return vertices_size(); }
  unsigned n_faces() const { return faces_size(); }

  Point2 position(int v) const { return {xy_[v], xy_[vertices_size() + v]}; }
  void set_position(int v, Point2 p) {
    xy_[v] = p.x;
    xy_[vertices_size() + v] = p.y;
  }
  std::array<int, 4> vertices(int f) const {
    const int i = f / n_, j = f % n_;
    return {j + i * (n_ + 1), j + (i + 1) * (n_ + 1), (j + 1) + (i + 1) * (n_ + 1), (j + 1) + i * (n_ + 1)};
  }
  std::vector<double>& xy() { return xy_; }
  const std::vector<double>& xy() const { return xy_; }

  // OFF writer, "%.10f" coordinates with z = 0 (otlib/surface_mesh/IO_off.cpp)
  bool write(const std::string& path) const {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "OFF\n%u %u 0\n", vertices_size(), faces_size());
    for (unsigned v = 0; v < vertices_size(); ++v)
      std::fprintf(f, "%.10f %.10f %.10f\n", xy_[v], xy_[vertices_size() + v], 0.0);
    for (unsigned q = 0; q < faces_size(); ++q) {
      const auto vs = vertices(int(q));
      std::fprintf(f, "4 %d %d %d %d\n", vs[0], vs[1], vs[2], vs[3]);
    }
    return std::fclose(f) == 0;
  }

 private:
  int n_ = 0;
  std::vector<double> xy_;
};

// otlib/transport_map.h:29-63.  Copies share the inverse-map BVH (the
// reference copies a raw BVH pointer and double-deletes it).
class TransportMap {
 public:
  TransportMap() = default;
  TransportMap(std::shared_ptr<GridMesh> origin, std::shared_ptr<GridMesh> fwd,
               std::shared_ptr<std::vector<double>> density, int device = 0)
      : origin_(std::move(origin)), fwd_(std::move(fwd)), density_(std::move(density)), device_(device) {}

  /** builds the BVH; also done lazily by the first inverse query */
  void init_inverse() const {
    if (!handle()) return;
    if (otm_map_init_inverse(handle()) != OTM_OK) std::cerr << "otmap init_inverse: " << otm_last_error() << "\n";
  }

  Point2 inv(Point2 p) const { return query(p, 0); }
  Point2 inv_fast(Point2 p) const { return query(p, 1); }

  /** batched inverse of interleaved points; mode 0 = inv, 1 = inv_fast */
  bool inverse(const double* pts, long count, double* out, int mode = 0, int* face = nullptr) const {
    if (!handle()) return false;
    return otm_map_inverse(handle(), pts, count, mode, out, face, nullptr) == OTM_OK;
  }

  std::shared_ptr<GridMesh> fwd_mesh_ptr() { return fwd_; }
  std::shared_ptr<std::vector<double>> density_ptr() { return density_; }
  const GridMesh& origin_mesh() const { return *origin_; }
  const GridMesh& fwd_mesh() const { return *fwd_; }
  const std::vector<double>& density() const { return *density_; }
  bool valid() const { return fwd_ != nullptr; }

 private:
  struct Handle {
    otm_map* m = nullptr;
    ~Handle() { otm_map_destroy(m); }
  };
  otm_map* handle() const {
    if (!fwd_) return nullptr;
    if (!bvh_) {
      auto h = std::make_shared<Handle>();
      if (otm_map_create(&h->m, device_, fwd_->grid_size(), fwd_->xy().data(), density_->data()) != OTM_OK) {
        std::cerr << "otmap TransportMap: " << otm_last_error() << "\n";
        return nullptr;
      }
      bvh_ = h;
    }
    return bvh_->m;
  }
  Point2 query(Point2 p, int mode) const {
    double in[2] = {p.x, p.y}, out[2] = {0., 0.};
    inverse(in, 1, out, mode);
    return {out[0], out[1]};
  }

  std::shared_ptr<GridMesh> origin_, fwd_;
  std::shared_ptr<std::vector<double>> density_;
  int device_ = 0;
  mutable std::shared_ptr<Handle> bvh_;
};

// otlib/transport_map.cpp:105-128: points are replaced by their inverse images
inline void apply_inverse_map(const TransportMap& tmap, std::vector<Point2>& points, int verbose_level = 2) {
  (void)verbose_level;
  std::vector<double> in(2 * points.size()), out(2 * points.size());
  for (size_t k = 0; k < points.size(); ++k) {
    in[2 * k] = points[k].x;
    in[2 * k + 1] = points[k].y;
  }
  tmap.init_inverse();
  if (!tmap.inverse(in.data(), long(points.size()), out.data(), 0)) return;
  for (size_t k = 0; k < points.size(); ++k) points[k] = {out[2 * k], out[2 * k + 1]};
}

}  // namespace otmap
