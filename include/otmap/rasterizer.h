// Drop-in for the reference's otlib/utils/rasterizer.h (ggael/otmap):
// rasterize_image in per-face density mode.  The image is row-major,
// img[y * cols + x] == the reference's img(y, x).  Header-only over the C ABI
// of libotmap_amd.so.
#pragma once

#include <iostream>
#include <vector>

#include "../otmap_capi.h"
#include "transport_map.h"

namespace otmap {

enum RasterImageOption { RIO_PerVertexDensity, RIO_PerFaceDensity };

// otlib/utils/rasterizer.h:21 (density per face; RIO_PerFaceDensity only)
inline bool rasterize_image(const GridMesh& mesh, const std::vector<double>& density_per_face, int rows, int cols,
                            std::vector<double>& img, int device = 0) {
  img.assign(size_t(rows) * cols, 0.);
  if (otm_rasterize(device, mesh.grid_size(), mesh.xy().data(), density_per_face.data(), rows, cols, img.data()) !=
      OTM_OK) {
    std::cerr << "otmap rasterize_image: " << otm_last_error() << "\n";
    return false;
  }
  return true;
}

// otlib/utils/rasterizer.h:22 (unit density per face)
inline bool rasterize_image(const GridMesh& mesh, int rows, int cols, std::vector<double>& img, int device = 0) {
  img.assign(size_t(rows) * cols, 0.);
  if (otm_rasterize(device, mesh.grid_size(), mesh.xy().data(), nullptr, rows, cols, img.data()) != OTM_OK) {
    std::cerr << "otmap rasterize_image: " << otm_last_error() << "\n";
    return false;
  }
  return true;
}

}  // namespace otmap
