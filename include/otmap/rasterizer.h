// Drop-in for the reference's otlib/utils/rasterizer.h (ggael/otmap):
// rasterize_image into an image of the caller's size (img.rows() x
// img.cols(); Eigen::MatrixXd when Eigen is on the include path), per-face
// density mode.  Header-only over the C ABI of libotmap_amd.so.
#pragma once

#include <iostream>
#include <vector>

#include "../otmap_capi.h"
#include "transport_map.h"

namespace otmap {

// otlib/utils/rasterizer.h:16-19
enum RasterImageOption { RIO_PerVertexDensity, RIO_PerFaceDensity };

namespace detail {
inline void rasterize(const GridMesh& mesh, const double* density, Image& img, RasterImageOption opt) {
  const long rows = img.rows(), cols = img.cols();
  for (long x = 0; x < cols; ++x)
    for (long y = 0; y < rows; ++y) img(y, x) = 0.;
  if (opt != RIO_PerFaceDensity) {
    std::cerr << "otmap rasterize_image: only RIO_PerFaceDensity is supported\n";
    return;
  }
  const std::vector<double> xy = mesh.xy();
  std::vector<double> rm(size_t(rows) * size_t(cols));  // row-major, as the C ABI writes it
  if (otm_rasterize(0, mesh.grid_size(), xy.data(), density, int(rows), int(cols), rm.data()) != OTM_OK) {
    report("rasterize_image");
    return;
  }
  for (long x = 0; x < cols; ++x)
    for (long y = 0; y < rows; ++y) img(y, x) = rm[size_t(y) * size_t(cols) + size_t(x)];
}
}  // namespace detail

// otlib/utils/rasterizer.h:21: every pixel covered by face f gets
// float(1/nf) * density(f) / |area(f)|, the largest covering face wins
inline void rasterize_image(const GridMesh& mesh, const DensityVector& density_per_face, Image& img,
                            RasterImageOption opt = RIO_PerFaceDensity) {
  if (size_t(density_per_face.size()) != size_t(mesh.faces_size())) {
    std::cerr << "otmap rasterize_image: density size does not match the mesh\n";
    return;
  }
  detail::rasterize(mesh, density_per_face.data(), img, opt);
}

// otlib/utils/rasterizer.h:22 (unit density per face)
inline void rasterize_image(const GridMesh& mesh, Image& img, RasterImageOption opt = RIO_PerFaceDensity) {
  detail::rasterize(mesh, nullptr, img, opt);
}

}  // namespace otmap
