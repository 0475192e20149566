// Point-in-quad predicates shared by the inverse-map and raster kernels and the
// host BVH build.  Operation order follows the reference so that results are
// bit-identical (the library is built with -ffp-contract=off):
//   cross2 / signed_area          otlib/utils/eigen_addons.h:16-23
//   triangle coordinates          otlib/utils/mesh_utils.h:18-25
//   inside_quad                   otlib/utils/mesh_utils.cpp:62-74
//   bilinear coords in a quad     otlib/utils/mesh_utils.cpp:18-59, 76-88
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace otm {
namespace geo {

struct P2 {
  double x, y;
};

__host__ __device__ inline double cross2(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }
__host__ __device__ inline P2 sub(P2 a, P2 b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ inline double sarea(P2 a, P2 b, P2 c) { return 0.5 * cross2(sub(b, a), sub(c, a)); }

__host__ __device__ inline void tri_coords(P2 q, P2 p0, P2 p1, P2 p2, double& u, double& v) {
  const P2 q2 = sub(q, p2), eu = sub(p1, p2), ev = sub(p0, p2);
  const double a2 = cross2(ev, eu);
  u = cross2(q2, eu) / a2;
  v = cross2(ev, q2) / a2;
}

// q inside one of the triangles (p0,p1,p2), (p0,p2,p3) up to 1e-8
__host__ __device__ inline bool inside_quad(P2 q, const P2* p) {
  const double eps = 1e-8;
  double u, v;
  tri_coords(q, p[0], p[1], p[2], u, v);
  if (u >= -eps && v >= -eps && u + v <= 1. + eps) return true;
  tri_coords(q, p[0], p[2], p[3], u, v);
  if (u >= -eps && v >= -eps && u + v <= 1. + eps) return true;
  return false;
}

// Inverse bilinear interpolation: weights of the 4 corners.  Returns false
// (leaving w untouched) when the quadratic has no real root (D < 0).
__host__ __device__ inline bool bilinear_weights(P2 q, const P2* p, double* w) {
  const double A0 = sarea(q, p[0], p[1]);
  const double A1 = sarea(q, p[1], p[2]);
  const double A2 = sarea(q, p[2], p[3]);
  const double A3 = sarea(q, p[3], p[0]);
  const double B0 = sarea(q, p[3], p[1]);
  const double B1 = sarea(q, p[0], p[2]);
  const double B3 = sarea(q, p[2], p[0]);
  const double D = B0 * B0 + B1 * B1 + 2 * A0 * A2 + 2 * A1 * A3;
  if (D < 0) return false;
  const double sD = sqrt(D);
  const double E0 = 2 * A0 - B0 - B1 + sD;
  const double E3 = 2 * A3 - B3 - B0 + sD;
  const double zero = 4 * DBL_MIN;
  double v = fabs(A0) <= zero ? 0 : 2 * A0 / E0;
  double u = fabs(A3) <= zero ? 0 : 2 * A3 / E3;
  const double E0p = 2 * A0 - B0 - B1 - sD;
  const double E3p = 2 * A3 - B3 - B0 - sD;
  double v_alt = fabs(A0) <= zero ? 0 : 2 * A0 / E0p;
  double u_alt = fabs(A3) <= zero ? 0 : 2 * A3 / E3p;
  const double eps = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON)
  if (fabs(u) < eps) u = 0;
  if (fabs(v) < eps) v = 0;
  if (fabs(u - 1.) < eps) u = 1;
  if (fabs(v - 1.) < eps) v = 1;
  if (fabs(u_alt) < eps) u_alt = 0;
  if (fabs(v_alt) < eps) v_alt = 0;
  if (fabs(u_alt - 1.) < eps) u_alt = 1;
  if (fabs(v_alt - 1.) < eps) v_alt = 1;
  if ((u < 0 || v < 0 || u > 1 || v > 1) && (u_alt >= 0 && v_alt >= 0 && u_alt <= 1 && v_alt <= 1)) {
    double t = u; u = u_alt; u_alt = t;
    t = v; v = v_alt; v_alt = t;
  }
  w[0] = (1. - u) * (1 - v);
  w[1] = u * (1. - v);
  w[2] = u * v;
  w[3] = (1. - u) * v;
  const double e4 = 4 * DBL_EPSILON;
  for (int k = 0; k < 4; ++k) {
    if (fabs(w[k]) <= e4) w[k] = 0;
    else if (fabs(w[k] - 1.) <= e4) w[k] = 1;
  }
  return true;
}

// all four weights in [0, 1] (compared against float literals, as the reference)
__host__ __device__ inline bool weights_in_unit(const double* w) {
  return w[0] <= 1.0f && w[1] <= 1.0f && w[2] <= 1.0f && w[3] <= 1.0f && w[0] >= 0.0f && w[1] >= 0.0f &&
         w[2] >= 0.0f && w[3] >= 0.0f;
}

}  // namespace geo
}  // namespace otm
