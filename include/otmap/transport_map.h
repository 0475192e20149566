// Drop-in for the reference's otlib/transport_map.h and the subset of
// surface_mesh::Surface_mesh its callers use on a regular quad grid
// (ggael/otmap).  Header-only over the C ABI of libotmap_amd.so.
//
// Types follow the reference when Eigen is on the include path (the
// reference's own build always has it): points are Eigen::Vector2d
// (surface_mesh/types.h:39-42), densities Eigen::VectorXd, images
// Eigen::MatrixXd.  Without Eigen the same names are small value types with
// the members the callers use (x(), y(), data(), size(), rows(), cols(),
// operator()), so a caller compiles against either.
//
// GridMesh: (n+1)^2 vertices, vertex (i, j) -> index j + i*(n+1); n^2 quads,
// quad (i, j) -> index j + i*n with vertices (i,j), (i+1,j), (i+1,j+1),
// (i,j+1) (otlib/utils/mesh_utils.cpp:90-122; the face circulator starts at
// the first vertex, otlib/surface_mesh/Surface_mesh.cpp:475).
#pragma once

#include <algorithm>
#include <array>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../otmap_capi.h"

#if defined(OTMAP_NO_EIGEN)
#define OTMAP_HAS_EIGEN 0
#elif __has_include(<Eigen/Core>)
#include <Eigen/Core>
#define OTMAP_HAS_EIGEN 1
#else
#define OTMAP_HAS_EIGEN 0
#endif

namespace otmap {

#if OTMAP_HAS_EIGEN
typedef Eigen::Vector2d Point;
typedef Eigen::VectorXd DensityVector;
typedef Eigen::MatrixXd Image;
#else
// 2-D point with the accessors of Eigen::Vector2d that callers use
struct Point {
  double v[2] = {0., 0.};
  Point() = default;
  Point(double x, double y) : v{x, y} {}
  double& x() { return v[0]; }
  double& y() { return v[1]; }
  double x() const { return v[0]; }
  double y() const { return v[1]; }
  double& operator[](int k) { return v[k]; }
  double operator[](int k) const { return v[k]; }
};
typedef std::vector<double> DensityVector;
// rows x cols FP64 image stored column-major, as Eigen::MatrixXd
class Image {
 public:
  Image() = default;
  Image(long rows, long cols) : r_(rows), c_(cols), d_(size_t(rows * cols), 0.) {}
  long rows() const { return r_; }
  long cols() const { return c_; }
  long size() const { return r_ * c_; }
  double& operator()(long y, long x) { return d_[size_t(x * r_ + y)]; }
  double operator()(long y, long x) const { return d_[size_t(x * r_ + y)]; }
  double* data() { return d_.data(); }
  const double* data() const { return d_.data(); }
  void setZero() { std::fill(d_.begin(), d_.end(), 0.); }

 private:
  long r_ = 0, c_ = 0;
  std::vector<double> d_;
};
#endif

class GridMesh {
 public:
  // handles of the reference's Surface_mesh (Surface_mesh.h:43-120)
  struct Vertex {
    explicit Vertex(int i = -1) : i_(i) {}
    int idx() const { return i_; }
    bool is_valid() const { return i_ >= 0; }
    bool operator==(const Vertex& o) const { return i_ == o.i_; }
    bool operator!=(const Vertex& o) const { return i_ != o.i_; }
    int i_;
  };
  struct Face {
    explicit Face(int i = -1) : i_(i) {}
    int idx() const { return i_; }
    bool is_valid() const { return i_ >= 0; }
    int i_;
  };
  template <class H>
  class Range {  // the handles 0 .. count-1
   public:
    struct iterator {
      int i;
      H operator*() const { return H(i); }
      iterator& operator++() { ++i; return *this; }
      bool operator!=(const iterator& o) const { return i != o.i; }
    };
    explicit Range(int count) : count_(count) {}
    iterator begin() const { return {0}; }
    iterator end() const { return {count_}; }

   private:
    int count_;
  };

  GridMesh() = default;
  // vertex positions SoA: x of all (n+1)^2 vertices, then y
  GridMesh(int n, const std::vector<double>& xy) : n_(n), pts_(size_t(n + 1) * (n + 1)) {
    const size_t V = pts_.size();
    for (size_t v = 0; v < V && 2 * V <= xy.size(); ++v) pts_[v] = Point(xy[v], xy[V + v]);
  }

  // generate_quad_mesh(n+1, n+1) (otlib/utils/mesh_utils.cpp:90-122):
  // vertex (i, j) at (i * (1/n), j * (1/n))
  static GridMesh regular(int n) {
    GridMesh m;
    m.n_ = n;
    m.pts_.resize(size_t(n + 1) * (n + 1));
    const double dx = 1. / double(n);
    for (int i = 0; i <= n; ++i)
      for (int j = 0; j <= n; ++j) m.pts_[j + size_t(i) * (n + 1)] = Point(double(i) * dx, double(j) * dx);
    return m;
  }

  int grid_size() const { return n_; }
  unsigned vertices_size() const { return unsigned(pts_.size()); }
  unsigned faces_size() const { return unsigned(n_ * n_); }
  unsigned n_vertices() const { return vertices_size(); }
  unsigned n_faces() const { return faces_size(); }
  bool empty() const { return pts_.empty(); }
  void clear() {
    n_ = 0;
    pts_.clear();
  }

  std::vector<Point>& points() { return pts_; }
  const std::vector<Point>& points() const { return pts_; }
  Point& position(Vertex v) { return pts_[size_t(v.idx())]; }
  const Point& position(Vertex v) const { return pts_[size_t(v.idx())]; }
  Range<Vertex> vertices() const { return Range<Vertex>(int(vertices_size())); }
  Range<Face> faces() const { return Range<Face>(int(faces_size())); }
  // the quad's vertices in circulator order
  std::array<Vertex, 4> vertices(Face f) const {
    const int i = f.idx() / n_, j = f.idx() % n_;
    return {Vertex(j + i * (n_ + 1)), Vertex(j + (i + 1) * (n_ + 1)), Vertex((j + 1) + (i + 1) * (n_ + 1)),
            Vertex((j + 1) + i * (n_ + 1))};
  }

  // positions as the C ABI takes them: SoA, x of every vertex, then y
  std::vector<double> xy() const {
    const size_t V = pts_.size();
    std::vector<double> out(2 * V);
    for (size_t v = 0; v < V; ++v) {
      out[v] = pts_[v].x();
      out[V + v] = pts_[v].y();
    }
    return out;
  }
  void set_xy(const double* xy) {
    const size_t V = pts_.size();
    for (size_t v = 0; v < V; ++v) pts_[v] = Point(xy[v], xy[V + v]);
  }

  // OFF writer, "%.10f" coordinates with z = 0 (otlib/surface_mesh/IO_off.cpp:312-344)
  bool write(const std::string& path) const {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "OFF\n%u %u 0\n", vertices_size(), faces_size());
    for (const Point& p : pts_) std::fprintf(f, "%.10f %.10f %.10f\n", p.x(), p.y(), 0.0);
    for (unsigned q = 0; q < faces_size(); ++q) {
      const auto vs = vertices(Face(int(q)));
      std::fprintf(f, "4 %d %d %d %d\n", vs[0].idx(), vs[1].idx(), vs[2].idx(), vs[3].idx());
    }
    return std::fclose(f) == 0;
  }

 private:
  int n_ = 0;
  std::vector<Point> pts_;
};

namespace detail {
inline void report(const char* what) { std::cerr << "otmap " << what << ": " << otm_last_error() << "\n"; }
inline void points_to_xy(const std::vector<Point>& pts, std::vector<double>& xy) {
  xy.resize(2 * pts.size());
  for (size_t k = 0; k < pts.size(); ++k) {
    xy[2 * k] = pts[k].x();
    xy[2 * k + 1] = pts[k].y();
  }
}
}  // namespace detail

// otlib/transport_map.h:29-63.  Copies share the inverse-map BVH (the
// reference copies a raw BVH pointer and double-deletes it).
class TransportMap {
 public:
  TransportMap() = default;
  TransportMap(std::shared_ptr<GridMesh> origin_mesh, std::shared_ptr<GridMesh> fwd_mesh,
               std::shared_ptr<DensityVector> density, int device = 0)
      : origin_(std::move(origin_mesh)), fwd_(std::move(fwd_mesh)), density_(std::move(density)), device_(device) {}

  /** this function must be called at least once before calling inv/inv_fast
   *  (the first inverse query also builds the BVH) */
  void init_inverse() const {
    if (otm_map* h = handle())
      if (otm_map_init_inverse(h) != OTM_OK) detail::report("init_inverse");
  }

  Point inv(const Point& p) const { return query(p, 0); }
  Point inv_fast(const Point& p) const { return query(p, 1); }

  /** batched inverse of `count` interleaved points; mode 0 = inv, 1 =
   *  inv_fast; face (nullable): selected face per point, -1 if none */
  bool inverse(const double* pts, long count, double* out, int mode = 0, int* face = nullptr) const {
    otm_map* h = handle();
    if (!h) return false;
    if (otm_map_inverse(h, pts, count, mode, out, face, nullptr) != OTM_OK) {
      detail::report("inverse");
      return false;
    }
    return true;
  }

  std::shared_ptr<GridMesh> fwd_mesh_ptr() { return fwd_; }
  std::shared_ptr<DensityVector> density_ptr() { return density_; }
  const GridMesh& origin_mesh() const { return *origin_; }
  const GridMesh& fwd_mesh() const { return *fwd_; }
  const DensityVector& density() const { return *density_; }
  bool valid() const { return fwd_ != nullptr; }

 private:
  struct Handle {
    otm_map* m = nullptr;
    ~Handle() { otm_map_destroy(m); }
  };
  otm_map* handle() const {
    if (!fwd_ || !density_) return nullptr;
    if (!bvh_) {
      auto h = std::make_shared<Handle>();
      const std::vector<double> xy = fwd_->xy();
      if (otm_map_create(&h->m, device_, fwd_->grid_size(), xy.data(), density_->data()) != OTM_OK) {
        detail::report("TransportMap");
        return nullptr;
      }
      bvh_ = h;
    }
    return bvh_->m;
  }
  Point query(const Point& p, int mode) const {
    double in[2] = {p.x(), p.y()}, out[2] = {0., 0.};
    inverse(in, 1, out, mode);
    return Point(out[0], out[1]);
  }

  std::shared_ptr<GridMesh> origin_, fwd_;
  std::shared_ptr<DensityVector> density_;
  int device_ = 0;
  mutable std::shared_ptr<Handle> bvh_;
};

/** Inverts uniform mesh relative to a transport map
 *  (otlib/transport_map.h:65-68, transport_map.cpp:105-128): every point is
 *  replaced by its inverse image; points no face holds become NaN. */
inline void apply_inverse_map(const TransportMap& tmap, std::vector<Point>& points, int verbose_level = 2) {
  (void)verbose_level;
  std::vector<double> in, out(2 * points.size());
  detail::points_to_xy(points, in);
  tmap.init_inverse();
  if (!tmap.inverse(in.data(), long(points.size()), out.data(), 0)) return;
  for (size_t k = 0; k < points.size(); ++k) points[k] = Point(out[2 * k], out[2 * k + 1]);
}

/** otlib/transport_map.h:70 (transport_map.cpp:131-215): transport cost
 *  between two meshes of the same grid, per-face density; cost_per_face
 *  (nullable) receives each face's normalised cost. */
inline double transport_cost(const GridMesh& src_mesh, const GridMesh& dst_mesh, const DensityVector& density_per_face,
                             DensityVector* cost_per_face = 0) {
  const int n = src_mesh.grid_size();
  if (dst_mesh.grid_size() != n || size_t(density_per_face.size()) != size_t(n) * n) {
    std::cerr << "otmap transport_cost: meshes and density of different grid sizes\n";
    return 0.;
  }
  const std::vector<double> s = src_mesh.xy(), d = dst_mesh.xy();
  std::vector<double> cost(cost_per_face ? size_t(n) * n : 0);
  double total = 0.;
  if (otm_transport_cost(0, n, s.data(), d.data(), density_per_face.data(), cost_per_face ? cost.data() : nullptr,
                         &total) != OTM_OK) {
    detail::report("transport_cost");
    return 0.;
  }
  if (cost_per_face) {
    cost_per_face->resize(long(cost.size()));
    for (size_t f = 0; f < cost.size(); ++f) (*cost_per_face)[long(f)] = cost[f];
  }
  return total;
}

}  // namespace otmap
