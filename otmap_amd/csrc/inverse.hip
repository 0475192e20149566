// TransportMap inverse: BVH over the forward quads + one GPU query per lane.
//
// Reference: TransportMap::init_inverse / inv / inv_fast / apply_inverse_map
// (otlib/transport_map.cpp:36-128) over BVH2D (otlib/utils/bvh2d.cpp:29-243).
//
// Build: on the device (bvh_build.hip), the reference's tree node for node:
// face centroids; per node the AABB of its faces' vertices inflated by
// 2*eps*diag + eps; leaf when <= 4 faces or depth >= 24; otherwise the longest
// axis split at the float-rounded box midpoint with the reference's in-place
// two-pointer partition.  The same partition keeps the same face order inside
// leaves and the same child order, so hits come out in the reference's
// traversal order.
//
// Device layout: nodes flattened to 48-byte records {box, first, count,
// escape}, where `escape` is the next node in depth-first order once a
// subtree is finished or skipped (child1 -> child2 -> parent's escape).  A
// query then walks the tree without a stack in exactly the reference's
// recursive order (child 1 before child 2).
//
// Selection in inv() with several hits keeps the reference's behaviour
// (otlib/transport_map.cpp:70-85): the face may switch to a later hit of
// higher density, the weights stay those of the first hit, and after the
// first switch the comparison value becomes the bool true (1.0).

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <vector>

#include "../../include/otmap_capi.h"
#include "geometry.h"
#include "otm_internal.h"

using namespace otm;
using namespace otm::geo;

namespace otm {

namespace {

__host__ __device__ inline void quad_vertices(int n, int f, int* v) {
  const int i = f / n, j = f - (f / n) * n;
  v[0] = j + i * (n + 1);
  v[1] = j + (i + 1) * (n + 1);
  v[2] = (j + 1) + (i + 1) * (n + 1);
  v[3] = (j + 1) + i * (n + 1);
}

__device__ inline bool box_contains(const BNode& b, P2 p) {
  return b.lox <= p.x && b.loy <= p.y && p.x <= b.hix && p.y <= b.hiy;
}

// Inverse image of one point q (transport_map.cpp:50-102).  fast: first hit
// (inv_fast), else every hit with the reference's selection.  Returns the
// number of hits; *face the selected face (-1 if none); (rx, ry) its
// bilinear interpolation of the origin positions (NaN if none).
__device__ inline int inv_query(const BNode* __restrict__ nodes, const int* __restrict__ faces,
                                const double* __restrict__ xy, const double* __restrict__ density, int n, P2 q,
                                int fast, double& rx, double& ry, int& face) {
  const long V = long(n + 1) * (n + 1);
  q.x = fmin(fmax(q.x, 0.), 1.);  // snap to [0,1]^2 (transport_map.cpp:53)
  q.y = fmin(fmax(q.y, 0.), 1.);
  int hits = 0;
  int chosen = -1;
  double w0[4] = {0., 0., 0., 0.};
  double best = 0.;
  // invariant: `node` holds q (or is -1 = traversal finished)
  int node = box_contains(nodes[0], q) ? 0 : -1;
  while (node >= 0) {
    const BNode nd = nodes[node];
    int next;
    if (nd.count < 0) {
      next = nd.first;  // inner node: child 1 first
    } else {
      bool stop = false;
      for (int s = nd.first; s < nd.first + nd.count; ++s) {
        const int f = faces[s];
        int v[4];
        quad_vertices(n, f, v);
        P2 p[4];
  #pragma unroll
        for (int t = 0; t < 4; ++t) p[t] = {xy[v[t]], xy[V + v[t]]};
        if (!inside_quad(q, p)) continue;
        double w[4];
        if (!bilinear_weights(q, p, w)) continue;
        if (!weights_in_unit(w)) continue;
        if (hits == 0) {
          chosen = f;
          for (int t = 0; t < 4; ++t) w0[t] = w[t];
          best = density[f];
        } else if (density[f] > best) {
          chosen = f;
          best = (density[f] > best) ? 1.0 : 0.0;
        }
        ++hits;
        if (fast) { stop = true; break; }
      }
      if (stop) break;
      next = nd.escape;  // leaf done: next subtree in depth-first order
    }
    // skip subtrees whose box does not hold q (child1 -> child2 -> parent's escape ...)
    while (next >= 0 && !box_contains(nodes[next], q)) next = nodes[next].escape;
    node = next;
  }
  face = chosen;
  if (hits == 0) {
    rx = __builtin_nan("");
    ry = __builtin_nan("");
    return 0;
  }
  int v[4];
  quad_vertices(n, chosen, v);
  const double dx = 1. / double(n);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int vi = v[t] / (n + 1), vj = v[t] - vi * (n + 1);
    const double ox = double(vi) * dx, oy = double(vj) * dx;
    if (t == 0) { rx = w0[0] * ox; ry = w0[0] * oy; }
    else { rx += w0[t] * ox; ry += w0[t] * oy; }
  }
  return hits;
}

// One query per lane over interleaved points.
__global__ __launch_bounds__(256) void k_inverse(const BNode* __restrict__ nodes, const int* __restrict__ faces,
                                                const double* __restrict__ xy, const double* __restrict__ density,
                                                int n, const double* __restrict__ pts, long m, int fast,
                                                double* __restrict__ out, int* __restrict__ face_out,
                                                int* __restrict__ nhits_out) {
  const long k = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= m) return;
  double rx, ry;
  int face;
  const int hits = inv_query(nodes, faces, xy, density, n, P2{pts[2 * k], pts[2 * k + 1]}, fast, rx, ry, face);
  if (nhits_out) nhits_out[k] = hits;
  if (face_out) face_out[k] = face;
  out[2 * k] = rx;
  out[2 * k + 1] = ry;
}

// Inverse images of the origin grid vertices (i/n, j/n) in vertex order, SoA
// output: the inverse map as a mesh (apps/barycenters.cpp:187-190).
__global__ __launch_bounds__(256) void k_inverse_vertices(const BNode* __restrict__ nodes,
                                                         const int* __restrict__ faces,
                                                         const double* __restrict__ xy,
                                                         const double* __restrict__ density, int n, int fast,
                                                         double* __restrict__ out) {
  const long V = long(n + 1) * (n + 1);
  const long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const int vi = int(v / (n + 1)), vj = int(v - long(vi) * (n + 1));
  const double dx = 1. / double(n);
  double rx, ry;
  int face;
  inv_query(nodes, faces, xy, density, n, P2{double(vi) * dx, double(vj) * dx}, fast, rx, ry, face);
  out[v] = rx;
  out[V + v] = ry;
}

}  // namespace
}  // namespace otm

struct otm_map {
  int device = 0;
  int n = 0;
  hipStream_t stream = nullptr;
  double* xy = nullptr;       // forward positions (device, 2V)
  double* density = nullptr;  // adjusted density (device, F)
  BNode* nodes = nullptr;
  int* faces = nullptr;
  int num_nodes = 0;
  double* d_pts = nullptr;    // staging for host queries
  double* d_out = nullptr;
  int* d_face = nullptr;
  int* d_nhits = nullptr;
  long staged = 0;
  double ms_build = 0.;
};

namespace {

int map_fail(const HipError& e) {
  set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
  return OTM_ERR_HIP;
}

void map_release_staging(otm_map* m) {
  if (m->d_pts) (void)hipFree(m->d_pts);
  if (m->d_out) (void)hipFree(m->d_out);
  if (m->d_face) (void)hipFree(m->d_face);
  if (m->d_nhits) (void)hipFree(m->d_nhits);
  m->d_pts = m->d_out = nullptr;
  m->d_face = m->d_nhits = nullptr;
  m->staged = 0;
}

// init_inverse(): BVH2D::build(fwd, 4, 24) on the device (otlib/transport_map.cpp:36-43)
void map_build(otm_map* m) {
  if (m->nodes) return;
  bvh_build_device(m->n, m->xy, 4, 24, m->stream, &m->nodes, &m->faces, &m->num_nodes);
}

void launch_inverse(otm_map* m, const double* d_pts, long count, int mode, double* d_out, int* d_face, int* d_nh) {
  if (count <= 0) return;
  const long blocks = (count + 255) / 256;
  hipLaunchKernelGGL(k_inverse, dim3(unsigned(blocks)), dim3(256), 0, m->stream, (const BNode*)m->nodes,
                     (const int*)m->faces, (const double*)m->xy, (const double*)m->density, m->n, d_pts, count,
                     mode, d_out, d_face, d_nh);
  OTM_HIP(hipGetLastError());
}

}  // namespace

extern "C" {

int otm_map_create(otm_map** out, int device, int n, const double* fwd_xy, const double* density) {
  if (!out || !fwd_xy || !density || n < 1) { set_error("invalid map arguments"); return OTM_ERR_ARG; }
  *out = nullptr;
  try {
    OTM_HIP(hipSetDevice(device));
    auto* m = new otm_map;
    m->device = device;
    m->n = n;
    const size_t V = size_t(n + 1) * (n + 1), F = size_t(n) * n;
    OTM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    OTM_HIP(hipMalloc(&m->xy, 2 * V * sizeof(double)));
    OTM_HIP(hipMalloc(&m->density, F * sizeof(double)));
    OTM_HIP(hipMemcpy(m->xy, fwd_xy, 2 * V * sizeof(double), hipMemcpyHostToDevice));
    OTM_HIP(hipMemcpy(m->density, density, F * sizeof(double), hipMemcpyHostToDevice));
    *out = m;
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

int otm_map_create_device(otm_map** out, int device, int n, const double* d_fwd_xy, const double* d_density) {
  if (!out || !d_fwd_xy || !d_density || n < 1) { set_error("invalid map arguments"); return OTM_ERR_ARG; }
  *out = nullptr;
  try {
    OTM_HIP(hipSetDevice(device));
    const size_t V = size_t(n + 1) * (n + 1), F = size_t(n) * n;
    auto* m = new otm_map;
    m->device = device;
    m->n = n;
    try {
      OTM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
      OTM_HIP(hipMalloc(&m->xy, 2 * V * sizeof(double)));
      OTM_HIP(hipMalloc(&m->density, F * sizeof(double)));
      OTM_HIP(hipMemcpyAsync(m->xy, d_fwd_xy, 2 * V * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
      OTM_HIP(hipMemcpyAsync(m->density, d_density, F * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
      OTM_HIP(hipStreamSynchronize(m->stream));
    } catch (...) {
      otm_map_destroy(m);
      throw;
    }
    *out = m;
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

int otm_map_destroy(otm_map* m) {
  if (!m) return OTM_OK;
  try {
    OTM_HIP(hipSetDevice(m->device));
    OTM_HIP(hipStreamSynchronize(m->stream));
    map_release_staging(m);
    if (m->nodes) (void)hipFree(m->nodes);
    if (m->faces) (void)hipFree(m->faces);
    if (m->xy) (void)hipFree(m->xy);
    if (m->density) (void)hipFree(m->density);
    (void)hipStreamDestroy(m->stream);
    delete m;
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

int otm_map_init_inverse(otm_map* m) {
  if (!m) { set_error("null map"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(m->device));
    map_build(m);
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

int otm_map_bvh_size(const otm_map* m) { return m ? m->num_nodes : 0; }

int otm_map_bvh_export(otm_map* m, double* boxes, int32_t* first, int32_t* count, int32_t* escape, int32_t* faces) {
  if (!m || !boxes || !first || !count || !escape || !faces) { set_error("invalid export arguments"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(m->device));
    map_build(m);
    std::vector<BNode> nodes(size_t(m->num_nodes));
    OTM_HIP(hipMemcpy(nodes.data(), m->nodes, nodes.size() * sizeof(BNode), hipMemcpyDeviceToHost));
    OTM_HIP(hipMemcpy(faces, m->faces, size_t(m->n) * m->n * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < nodes.size(); ++k) {
      boxes[4 * k + 0] = nodes[k].lox;
      boxes[4 * k + 1] = nodes[k].loy;
      boxes[4 * k + 2] = nodes[k].hix;
      boxes[4 * k + 3] = nodes[k].hiy;
      first[k] = nodes[k].first;
      count[k] = nodes[k].count;
      escape[k] = nodes[k].escape;
    }
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

int otm_map_inverse(otm_map* m, const double* pts, int64_t count, int mode, double* out, int32_t* face,
                    int32_t* nhits) {
  if (!m || (count > 0 && (!pts || !out)) || count < 0 || (mode != 0 && mode != 1)) {
    set_error("invalid inverse arguments");
    return OTM_ERR_ARG;
  }
  if (count == 0) return OTM_OK;
  try {
    OTM_HIP(hipSetDevice(m->device));
    map_build(m);
    if (count > m->staged) {
      map_release_staging(m);
      OTM_HIP(hipMalloc(&m->d_pts, 2 * size_t(count) * sizeof(double)));
      OTM_HIP(hipMalloc(&m->d_out, 2 * size_t(count) * sizeof(double)));
      OTM_HIP(hipMalloc(&m->d_face, size_t(count) * sizeof(int)));
      OTM_HIP(hipMalloc(&m->d_nhits, size_t(count) * sizeof(int)));
      m->staged = count;
    }
    OTM_HIP(hipMemcpyAsync(m->d_pts, pts, 2 * size_t(count) * sizeof(double), hipMemcpyHostToDevice, m->stream));
    launch_inverse(m, m->d_pts, count, mode, m->d_out, m->d_face, m->d_nhits);
    OTM_HIP(hipMemcpyAsync(out, m->d_out, 2 * size_t(count) * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    if (face) OTM_HIP(hipMemcpyAsync(face, m->d_face, size_t(count) * sizeof(int), hipMemcpyDeviceToHost, m->stream));
    if (nhits)
      OTM_HIP(hipMemcpyAsync(nhits, m->d_nhits, size_t(count) * sizeof(int), hipMemcpyDeviceToHost, m->stream));
    OTM_HIP(hipStreamSynchronize(m->stream));
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

int otm_map_inverse_device(otm_map* m, const double* d_pts, int64_t count, int mode, double* d_out, int32_t* d_face,
                           int32_t* d_nhits) {
  if (!m || (count > 0 && (!d_pts || !d_out)) || count < 0 || (mode != 0 && mode != 1)) {
    set_error("invalid inverse arguments");
    return OTM_ERR_ARG;
  }
  if (count == 0) return OTM_OK;
  try {
    OTM_HIP(hipSetDevice(m->device));
    map_build(m);
    launch_inverse(m, d_pts, count, mode, d_out, d_face, d_nhits);
    OTM_HIP(hipStreamSynchronize(m->stream));
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}


int otm_map_inverse_vertices_device(otm_map* m, int mode, double* d_xy) {
  if (!m || !d_xy || (mode != 0 && mode != 1)) {
    set_error("invalid inverse arguments");
    return OTM_ERR_ARG;
  }
  try {
    OTM_HIP(hipSetDevice(m->device));
    map_build(m);
    const long V = long(m->n + 1) * (m->n + 1);
    hipLaunchKernelGGL(k_inverse_vertices, dim3(unsigned((V + 255) / 256)), dim3(256), 0, m->stream,
                       (const BNode*)m->nodes, (const int*)m->faces, (const double*)m->xy, (const double*)m->density,
                       m->n, mode, d_xy);
    OTM_HIP(hipGetLastError());
    OTM_HIP(hipStreamSynchronize(m->stream));
    return OTM_OK;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

int otm_map_inverse_vertices(otm_map* m, int mode, double* xy) {
  if (!m || !xy) {
    set_error("invalid inverse arguments");
    return OTM_ERR_ARG;
  }
  try {
    OTM_HIP(hipSetDevice(m->device));
    const long V = long(m->n + 1) * (m->n + 1);
    double* d = nullptr;
    OTM_HIP(hipMalloc(&d, 2 * V * sizeof(double)));
    int rc = otm_map_inverse_vertices_device(m, mode, d);
    if (rc == OTM_OK) {
      const hipError_t e = hipMemcpy(xy, d, 2 * V * sizeof(double), hipMemcpyDeviceToHost);
      (void)hipFree(d);
      OTM_HIP(e);
    } else {
      (void)hipFree(d);
    }
    return rc;
  } catch (const HipError& e) {
    return map_fail(e);
  }
}

}  // extern "C"
