// Kernels around the solver: transport costs of a map, interpolation of
// inverse-map meshes, mean normalisation of rasterised images and the
// stippling point prune.  They serve the reference's map-level functions and
// its two pipelines:
//   transport_cost                 otlib/transport_map.cpp:131-215
//   compute_transport_cost         otlib/otsolver_2dgrid.cpp:608-640
//   barycentres (lerp / bilerp of inverse meshes, mean-normalised rasters)
//                                  apps/barycenters.cpp:128-140,187-290
//   stippling (prune of advected points on zero density)
//                                  apps/stippling.cpp:184-197
// Per-face values keep the reference's operation order (built without FMA
// contraction); sums are deterministic two-pass reductions.

#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "geometry.h"
#include "otm_internal.h"
#include "reduce.h"

namespace otm {
namespace {

constexpr int RB = 1024;  // max workgroups of the reductions below

using geo::P2;
using geo::sarea;

__device__ inline int vtx(int n, int i, int j) { return j + i * (n + 1); }

// vertices (i,j), (i+1,j), (i+1,j+1), (i,j+1) of face f = j + i n
__device__ inline void quad_v(int n, int f, int* v) {
  const int i = f / n, j = f - i * n;
  v[0] = vtx(n, i, j);
  v[1] = vtx(n, i + 1, j);
  v[2] = vtx(n, i + 1, j + 1);
  v[3] = vtx(n, i, j + 1);
}

// transport_cost on a quad grid mesh: the reference's per-face branch for
// nfaces == density size takes the triangle (from, to, next-to) of the face's
// first halfedge, i.e. quad vertices (3, 0, 1), and averages |displacement|^2
// at its three edge midpoints.
__global__ __launch_bounds__(NT) void k_tc_faces(int n, const double* __restrict__ src, const double* __restrict__ dst,
                                                const double* __restrict__ density, double* __restrict__ cost,
                                                double* partials) {
  const long V = long(n + 1) * (n + 1), F = long(n) * n;
  double acc[2] = {0., 0.};  // sum of costs, sum of density * area
  for (long f = long(blockIdx.x) * NT + threadIdx.x; f < F; f += long(gridDim.x) * NT) {
    int v[4];
    quad_v(n, int(f), v);
    const int a = v[3], b = v[0], c = v[1];
    const P2 sa = {src[a], src[V + a]}, sb = {src[b], src[V + b]}, sc = {src[c], src[V + c]};
    const double area = sarea(sa, sb, sc);
    const double dax = dst[a] - sa.x, day = dst[V + a] - sa.y;
    const double dbx = dst[b] - sb.x, dby = dst[V + b] - sb.y;
    const double dcx = dst[c] - sc.x, dcy = dst[V + c] - sc.y;
    auto d2 = [&](double l1, double l2, double l3) {
      const double x = l1 * dax + l2 * dbx + l3 * dcx;
      const double y = l1 * day + l2 * dby + l3 * dcy;
      return x * x + y * y;
    };
    double fc = (d2(0.5, 0.5, 0) + d2(0, 0.5, 0.5) + d2(0.5, 0, 0.5)) / 3.0;
    acc[1] += density[f] * area;
    fc = density[f] * area * fc;
    acc[0] += fc;
    if (cost) cost[f] = fc;
  }
  write_partials<2>(acc, partials);
}

// compute_transport_cost: 2 x 2 Gauss points of |interpolated displacement|^2
// per face, times density * A
__global__ __launch_bounds__(NT) void k_stc_faces(int n, const double* __restrict__ G, const double* __restrict__ rho,
                                                 double* __restrict__ cost, double* partials) {
  const long V = long(n + 1) * (n + 1), F = long(n) * n;
  const double A = 1.0 / (double(n) * double(n));
  const double z1 = sqrt(1. / 3.) / 2. + 0.5, z2 = -sqrt(1. / 3.) / 2. + 0.5;
  double acc[1] = {0.};
  for (long f = long(blockIdx.x) * NT + threadIdx.x; f < F; f += long(gridDim.x) * NT) {
    const int i = int(f / n), j = int(f - long(i) * n);
    const int v1 = vtx(n, i, j), v2 = vtx(n, i + 1, j), v3 = vtx(n, i + 1, j + 1), v4 = vtx(n, i, j + 1);
    auto d2 = [&](double u, double v) {
      const double x = (1 - u) * (1 - v) * G[v1] + u * (1 - v) * G[v2] + u * v * G[v3] + (1 - u) * v * G[v4];
      const double y =
          (1 - u) * (1 - v) * G[V + v1] + u * (1 - v) * G[V + v2] + u * v * G[V + v3] + (1 - u) * v * G[V + v4];
      return x * x + y * y;
    };
    const double c = rho[f] * A * (d2(z1, z1) + d2(z1, z2) + d2(z2, z2) + d2(z2, z1)) / 4.;
    if (cost) cost[f] = c;
    acc[0] += c;
  }
  write_partials<1>(acc, partials);
}

template <int K>
__global__ __launch_bounds__(NT) void k_fold(const double* partials, int nb, double* out) {
  double v[K];
  fold_partials<K>(v, partials, nb);
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) out[k] = v[k];
}

__global__ __launch_bounds__(NT) void k_div(double* __restrict__ x, long count, const double* __restrict__ d) {
  const double s = d[0];
  for (long i = long(blockIdx.x) * NT + threadIdx.x; i < count; i += long(gridDim.x) * NT) x[i] /= s;
}

// lerp / bilerp of K meshes (SoA positions), apps/barycenters.cpp:128-140
template <int K>
__global__ __launch_bounds__(NT) void k_mesh_lerp(long count, const double* __restrict__ m0,
                                                 const double* __restrict__ m1, const double* __restrict__ m2,
                                                 const double* __restrict__ m3, double u, double v,
                                                 double* __restrict__ out) {
  for (long i = long(blockIdx.x) * NT + threadIdx.x; i < count; i += long(gridDim.x) * NT) {
    if (K == 2) {
      out[i] = (1. - u) * m0[i] + u * m1[i];
    } else {
      out[i] = (1. - u) * (1. - v) * m0[i] + (u) * (1. - v) * m1[i] + (u) * (v)*m2[i] + (1. - u) * (v)*m3[i];
    }
  }
}

__global__ __launch_bounds__(NT) void k_img_sum(const double* __restrict__ img, long count, double* partials) {
  double acc[1] = {0.};
  for (long i = long(blockIdx.x) * NT + threadIdx.x; i < count; i += long(gridDim.x) * NT) acc[0] += img[i];
  write_partials<1>(acc, partials);
}

// img * (expected / mean), written as double or float
__global__ __launch_bounds__(NT) void k_img_scale(const double* __restrict__ img, long count,
                                                 const double* __restrict__ sum, double expected, int f32,
                                                 void* __restrict__ out) {
  const double scale = expected / (sum[0] / double(count));
  for (long i = long(blockIdx.x) * NT + threadIdx.x; i < count; i += long(gridDim.x) * NT) {
    const double v = img[i] * scale;
    if (f32) static_cast<float*>(out)[i] = float(v);
    else static_cast<double*>(out)[i] = v;
  }
}

// stippling prune (apps/stippling.cpp:188-196): keep point p when the density
// of its pixel (floor(p * n), row = y, column = x, face order y + x * n) is
// not zero; order preserved.  Pass 1: flags and per-workgroup counts.
constexpr int PRUNE_CHUNK = NT * 8;  // points per workgroup
// the prune predicate of point k (pixel floor(p n), clamped; NaN points dropped)
__device__ inline bool prune_keep(int n, const double* __restrict__ density, const double* __restrict__ pts, long k) {
  const double x = pts[2 * k], y = pts[2 * k + 1];
  int ix = int(floor(x * double(n))), iy = int(floor(y * double(n)));
  ix = ix < 0 ? 0 : (ix > n - 1 ? n - 1 : ix);
  iy = iy < 0 ? 0 : (iy > n - 1 ? n - 1 : iy);
  return !isnan(x) && !isnan(y) && density[long(iy) + long(ix) * n] != 0.;
}

__global__ __launch_bounds__(NT) void k_prune_count(int n, const double* __restrict__ density,
                                                   const double* __restrict__ pts, long count,
                                                   unsigned char* __restrict__ keep, int* __restrict__ block_count) {
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  int c = 0;
  const long base = long(blockIdx.x) * PRUNE_CHUNK;
  for (int u = threadIdx.x; u < PRUNE_CHUNK; u += NT) {
    const long k = base + u;
    if (k >= count) break;
    // the reference reads density(floor(y n), floor(x n)); indices are clamped
    // to the grid (a point at exactly 1.0 would read past the image there)
    const bool k_ok = prune_keep(n, density, pts, k);
    if (keep) keep[k] = k_ok ? 1 : 0;
    c += k_ok ? 1 : 0;
  }
  if (c) atomicAdd(&total, c);
  __syncthreads();
  if (threadIdx.x == 0) block_count[blockIdx.x] = total;
}

// pass 2 (one workgroup): exclusive scan of the block counts, total kept
__global__ __launch_bounds__(NT) void k_prune_scan(int* block_count, int nb, long* kept) {
  if (threadIdx.x != 0) return;
  long acc = 0;
  for (int b = 0; b < nb; ++b) {
    const int c = block_count[b];
    block_count[b] = int(acc);
    acc += c;
  }
  kept[0] = acc;
}

// pass 3: ordered scatter (each workgroup writes its chunk's kept points in
// order); keep == nullptr: the predicate is evaluated again instead of read
__global__ __launch_bounds__(NT) void k_prune_scatter(const double* __restrict__ pts, long count,
                                                     const unsigned char* __restrict__ keep,
                                                     const int* __restrict__ block_offset, double* __restrict__ out,
                                                     int n, const double* __restrict__ density) {
  __shared__ int off[NT + 1];
  const long base = long(blockIdx.x) * PRUNE_CHUNK;
  // thread t owns points [base + 8t, base + 8t + 8)
  constexpr int PER = PRUNE_CHUNK / NT;
  bool mine[PER];
  int c = 0;
  for (int u = 0; u < PER; ++u) {
    const long k = base + long(threadIdx.x) * PER + u;
    mine[u] = k < count && (keep ? keep[k] != 0 : prune_keep(n, density, pts, k));
    c += mine[u] ? 1 : 0;
  }
  off[threadIdx.x + 1] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    off[0] = 0;
    for (int t = 1; t <= NT; ++t) off[t] += off[t - 1];
  }
  __syncthreads();
  long dst = long(block_offset[blockIdx.x]) + off[threadIdx.x];
  for (int u = 0; u < PER; ++u) {
    const long k = base + long(threadIdx.x) * PER + u;
    if (mine[u]) {
      out[2 * dst] = pts[2 * k];
      out[2 * dst + 1] = pts[2 * k + 1];
      ++dst;
    }
  }
}

// Zero-allocation prune: the workgroup counts and the kept total live in a
// file-scope device array (one copy per GPU, held until the process ends)
// instead of buffers allocated per call.  Every call on a GPU shares that
// GPU's array: calls that overlap on one GPU race.  Nothing in it needs to
// start at zero (pass 1 writes every count it uses, pass 2 the total).
constexpr long PRUNE_MAX_BLOCKS = 8192;                           // workgroups of one call at most
constexpr long PRUNE_MAX_POINTS = PRUNE_MAX_BLOCKS * PRUNE_CHUNK;  // 16,777,216 points
static __device__ __align__(256) long long scratch[PRUNE_MAX_BLOCKS / 2 + 1];  // 8192 int counts, then the total

// generate_blue_noise_tile (apps/common/image_utils.cpp:148-173): the tile
// of nb points (N = floor(sqrt(nb)) per side) repeated over nt x nt cells
// of size N / n; candidate L = (i nt + j) nb + k is ((i + t_k.x) s, (j +
// t_k.y) s), kept when both coordinates are <= 1, in candidate order.
struct TileGrid {
  long nb;    // tile points
  int nt;     // cells per side
  double s;   // N / n
};
__device__ inline bool tile_point(const TileGrid g, const double* __restrict__ tile, long L, double& x, double& y) {
  const long cell = L / g.nb, k = L - cell * g.nb;
  const int i = int(cell / g.nt), j = int(cell - long(i) * g.nt);
  x = (double(i) + tile[2 * k]) * g.s;
  y = (double(j) + tile[2 * k + 1]) * g.s;
  return x <= 1. && y <= 1.;
}

__global__ __launch_bounds__(NT) void k_tile_count(const TileGrid g, const double* __restrict__ tile, long count,
                                                  int* __restrict__ block_count) {
  __shared__ int total;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  int c = 0;
  const long base = long(blockIdx.x) * PRUNE_CHUNK;
  for (int u = threadIdx.x; u < PRUNE_CHUNK; u += NT) {
    const long L = base + u;
    if (L >= count) break;
    double x, y;
    c += tile_point(g, tile, L, x, y) ? 1 : 0;
  }
  if (c) atomicAdd(&total, c);
  __syncthreads();
  if (threadIdx.x == 0) block_count[blockIdx.x] = total;
}

__global__ __launch_bounds__(NT) void k_tile_scatter(const TileGrid g, const double* __restrict__ tile, long count,
                                                    const int* __restrict__ block_offset, double* __restrict__ out) {
  __shared__ int off[NT + 1];
  const long base = long(blockIdx.x) * PRUNE_CHUNK;
  constexpr int PER = PRUNE_CHUNK / NT;
  double px[PER], py[PER];
  bool mine[PER];
  int c = 0;
  for (int u = 0; u < PER; ++u) {
    const long L = base + long(threadIdx.x) * PER + u;
    mine[u] = L < count && tile_point(g, tile, L, px[u], py[u]);
    c += mine[u] ? 1 : 0;
  }
  off[threadIdx.x + 1] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    off[0] = 0;
    for (int t = 1; t <= NT; ++t) off[t] += off[t - 1];
  }
  __syncthreads();
  long dst = long(block_offset[blockIdx.x]) + off[threadIdx.x];
  for (int u = 0; u < PER; ++u)
    if (mine[u]) {
      out[2 * dst] = px[u];
      out[2 * dst + 1] = py[u];
      ++dst;
    }
}

int red_blocks(long count) { return int(std::max<long>(1, std::min<long>((count + NT - 1) / NT, RB))); }

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { OTM_HIP(hipMalloc(&p, bytes)); }
  ~DevBuf() { if (p) (void)hipFree(p); }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

int fail(const HipError& e) {
  set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
  return OTM_ERR_HIP;
}

// cost (nullable) per face and the total of transport_cost on device buffers
void transport_cost_device(int n, const double* src, const double* dst, const double* density, double* cost,
                           double* h_total, hipStream_t s) {
  const long F = long(n) * n;
  const int nb = red_blocks(F);
  DevBuf part(2 * size_t(nb) * sizeof(double)), red(2 * sizeof(double));
  hipLaunchKernelGGL(k_tc_faces, dim3(nb), dim3(NT), 0, s, n, src, dst, density, cost, part.as<double>());
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_fold<2>), dim3(1), dim3(NT), 0, s, (const double*)part.as<double>(), nb, red.as<double>());
  OTM_HIP(hipGetLastError());
  if (cost) {  // cost_per_face /= integral
    hipLaunchKernelGGL(k_div, dim3(nb), dim3(NT), 0, s, cost, F, (const double*)(red.as<double>() + 1));
    OTM_HIP(hipGetLastError());
  }
  double hv[2];
  OTM_HIP(hipMemcpyAsync(hv, red.p, sizeof(hv), hipMemcpyDeviceToHost, s));
  OTM_HIP(hipStreamSynchronize(s));
  if (h_total) *h_total = hv[0] / hv[1];
}

void grid_transport_cost_device(int n, const double* G, const double* rho, double* cost, double* h_total,
                                hipStream_t s) {
  const long F = long(n) * n;
  const int nb = red_blocks(F);
  DevBuf part(size_t(nb) * sizeof(double)), red(sizeof(double));
  hipLaunchKernelGGL(k_stc_faces, dim3(nb), dim3(NT), 0, s, n, G, rho, cost, part.as<double>());
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_fold<1>), dim3(1), dim3(NT), 0, s, (const double*)part.as<double>(), nb, red.as<double>());
  OTM_HIP(hipGetLastError());
  double hv;
  OTM_HIP(hipMemcpyAsync(&hv, red.p, sizeof(hv), hipMemcpyDeviceToHost, s));
  OTM_HIP(hipStreamSynchronize(s));
  if (h_total) *h_total = hv;
}

// device address of this GPU's copy of `scratch` (looked up once per device,
// with that device current)
void* scratch_ptr(int device) {
  static std::mutex mu;
  static std::map<int, void*> ptrs;
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(device);
  if (it != ptrs.end()) return it->second;
  void* p = nullptr;
  OTM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  ptrs[device] = p;
  return p;
}

}  // namespace
}  // namespace otm

using namespace otm;

extern "C" {

int otm_transport_cost_device(int device, int n, const double* d_src_xy, const double* d_dst_xy,
                              const double* d_density, double* d_cost, double* total) {
  if (n < 1 || !d_src_xy || !d_dst_xy || !d_density) { set_error("invalid transport cost arguments"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(device));
    transport_cost_device(n, d_src_xy, d_dst_xy, d_density, d_cost, total, nullptr);
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

int otm_transport_cost(int device, int n, const double* src_xy, const double* dst_xy, const double* density,
                       double* cost, double* total) {
  if (n < 1 || !src_xy || !dst_xy || !density) { set_error("invalid transport cost arguments"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(device));
    const size_t V = size_t(n + 1) * (n + 1), F = size_t(n) * n;
    DevBuf s(2 * V * sizeof(double)), d(2 * V * sizeof(double)), rho(F * sizeof(double)), c(F * sizeof(double));
    OTM_HIP(hipMemcpy(s.p, src_xy, 2 * V * sizeof(double), hipMemcpyHostToDevice));
    OTM_HIP(hipMemcpy(d.p, dst_xy, 2 * V * sizeof(double), hipMemcpyHostToDevice));
    OTM_HIP(hipMemcpy(rho.p, density, F * sizeof(double), hipMemcpyHostToDevice));
    transport_cost_device(n, s.as<double>(), d.as<double>(), rho.as<double>(), cost ? c.as<double>() : nullptr,
                          total, nullptr);
    if (cost) OTM_HIP(hipMemcpy(cost, c.p, F * sizeof(double), hipMemcpyDeviceToHost));
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

int otm_grid_transport_cost(int device, int n, const double* disp_xy, const double* density, double* cost,
                            double* total) {
  if (n < 1 || !disp_xy || !density) { set_error("invalid transport cost arguments"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(device));
    const size_t V = size_t(n + 1) * (n + 1), F = size_t(n) * n;
    DevBuf g(2 * V * sizeof(double)), rho(F * sizeof(double)), c(F * sizeof(double));
    OTM_HIP(hipMemcpy(g.p, disp_xy, 2 * V * sizeof(double), hipMemcpyHostToDevice));
    OTM_HIP(hipMemcpy(rho.p, density, F * sizeof(double), hipMemcpyHostToDevice));
    grid_transport_cost_device(n, g.as<double>(), rho.as<double>(), cost ? c.as<double>() : nullptr, total, nullptr);
    if (cost) OTM_HIP(hipMemcpy(cost, c.p, F * sizeof(double), hipMemcpyDeviceToHost));
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

int otm_mesh_lerp_device(int device, long count, int k, const double* const* d_meshes, double u, double v,
                         double* d_out) {
  if (count < 1 || !d_meshes || !d_out || (k != 2 && k != 4)) { set_error("invalid lerp arguments"); return OTM_ERR_ARG; }
  for (int q = 0; q < k; ++q)
    if (!d_meshes[q]) { set_error("null mesh pointer"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(device));
    const int nb = int(std::min<long>((count + NT - 1) / NT, 4096));
    if (k == 2)
      hipLaunchKernelGGL((k_mesh_lerp<2>), dim3(nb), dim3(NT), 0, nullptr, count, d_meshes[0], d_meshes[1],
                         (const double*)nullptr, (const double*)nullptr, u, v, d_out);
    else
      hipLaunchKernelGGL((k_mesh_lerp<4>), dim3(nb), dim3(NT), 0, nullptr, count, d_meshes[0], d_meshes[1],
                         d_meshes[2], d_meshes[3], u, v, d_out);
    OTM_HIP(hipGetLastError());
    OTM_HIP(hipDeviceSynchronize());
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

int otm_image_normalize_mean_device(int device, long count, const double* d_img, double expected_mean, int out_f32,
                                    void* d_out, double* mean) {
  if (count < 1 || !d_img || !d_out) { set_error("invalid image arguments"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(device));
    const int nb = red_blocks(count);
    DevBuf part(size_t(nb) * sizeof(double)), red(sizeof(double));
    hipLaunchKernelGGL(k_img_sum, dim3(nb), dim3(NT), 0, nullptr, d_img, count, part.as<double>());
    OTM_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_fold<1>), dim3(1), dim3(NT), 0, nullptr, (const double*)part.as<double>(), nb,
                       red.as<double>());
    OTM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_img_scale, dim3(nb), dim3(NT), 0, nullptr, d_img, count, (const double*)red.as<double>(),
                       expected_mean, out_f32, d_out);
    OTM_HIP(hipGetLastError());
    double sum = 0.;
    OTM_HIP(hipMemcpy(&sum, red.p, sizeof(double), hipMemcpyDeviceToHost));
    if (mean) *mean = sum / double(count);
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

int otm_prune_points_device(int device, int n, const double* d_density, const double* d_pts, int64_t count,
                            double* d_out, int64_t* kept) {
  if (n < 1 || !d_density || !d_pts || !d_out || count < 0 || !kept) { set_error("invalid prune arguments"); return OTM_ERR_ARG; }
  if (count == 0) { *kept = 0; return OTM_OK; }
  try {
    OTM_HIP(hipSetDevice(device));
    const long nb = (count + PRUNE_CHUNK - 1) / PRUNE_CHUNK;
    DevBuf flags{size_t(count)}, blocks{size_t(nb) * sizeof(int)}, total{sizeof(long)};
    hipLaunchKernelGGL(k_prune_count, dim3(unsigned(nb)), dim3(NT), 0, nullptr, n, d_density, d_pts, long(count),
                       flags.as<unsigned char>(), blocks.as<int>());
    OTM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_prune_scan, dim3(1), dim3(NT), 0, nullptr, blocks.as<int>(), int(nb), total.as<long>());
    OTM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_prune_scatter, dim3(unsigned(nb)), dim3(NT), 0, nullptr, d_pts, long(count),
                       (const unsigned char*)flags.as<unsigned char>(), (const int*)blocks.as<int>(), d_out, n,
                       d_density);
    OTM_HIP(hipGetLastError());
    long k = 0;
    OTM_HIP(hipMemcpy(&k, total.p, sizeof(long), hipMemcpyDeviceToHost));
    *kept = k;
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

int otm_blue_noise_tile_device(int device, int n, const double* d_tile, int64_t nb, double* d_out, int64_t cap,
                               int64_t* count) {
  if (n < 1 || nb < 1 || !d_tile || !count || cap < 0 || (cap > 0 && !d_out)) {
    set_error("invalid blue-noise tile arguments");
    return OTM_ERR_ARG;
  }
  const int N = int(std::sqrt(double(nb)));
  const int nt = (n + N - 1) / N;
  const long total = long(nt) * nt * nb;
  if (total > (1L << 31) - PRUNE_CHUNK) { set_error("blue-noise tiling too large"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(device));
    const TileGrid g{long(nb), nt, double(N) / double(n)};
    const long nblk = (total + PRUNE_CHUNK - 1) / PRUNE_CHUNK;
    DevBuf blocks{size_t(nblk) * sizeof(int)}, kept{sizeof(long)};
    hipLaunchKernelGGL(k_tile_count, dim3(unsigned(nblk)), dim3(NT), 0, nullptr, g, d_tile, total, blocks.as<int>());
    OTM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_prune_scan, dim3(1), dim3(NT), 0, nullptr, blocks.as<int>(), int(nblk), kept.as<long>());
    OTM_HIP(hipGetLastError());
    long k = 0;
    OTM_HIP(hipMemcpy(&k, kept.p, sizeof(long), hipMemcpyDeviceToHost));
    *count = k;
    if (k > cap) {
      set_error("blue-noise tiling: output holds fewer points than the tiling produces (see *count)");
      return OTM_ERR_ARG;
    }
    hipLaunchKernelGGL(k_tile_scatter, dim3(unsigned(nblk)), dim3(NT), 0, nullptr, g, d_tile, total,
                       (const int*)blocks.as<int>(), d_out);
    OTM_HIP(hipGetLastError());
    OTM_HIP(hipDeviceSynchronize());
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

int otm_prune_points_device_noalloc(int device, int n, const double* d_density, const double* d_pts, int64_t count,
                                    double* d_out, int64_t* kept) {
  if (n < 1 || !d_density || !d_pts || !d_out || count < 0 || !kept) { set_error("invalid prune arguments"); return OTM_ERR_ARG; }
  if (count > PRUNE_MAX_POINTS) {
    set_error("otm_prune_points_device_noalloc: at most 16777216 points per call (its fixed device scratch); "
              "use otm_prune_points_device for more");
    return OTM_ERR_ARG;
  }
  if (count == 0) { *kept = 0; return OTM_OK; }
  try {
    OTM_HIP(hipSetDevice(device));
    auto* base = static_cast<int*>(scratch_ptr(device));
    int* blocks = base;
    long* total = reinterpret_cast<long*>(base + PRUNE_MAX_BLOCKS);
    const long nb = (count + PRUNE_CHUNK - 1) / PRUNE_CHUNK;
    hipLaunchKernelGGL(k_prune_count, dim3(unsigned(nb)), dim3(NT), 0, nullptr, n, d_density, d_pts, long(count),
                       (unsigned char*)nullptr, blocks);
    OTM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_prune_scan, dim3(1), dim3(NT), 0, nullptr, blocks, int(nb), total);
    OTM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_prune_scatter, dim3(unsigned(nb)), dim3(NT), 0, nullptr, d_pts, long(count),
                       (const unsigned char*)nullptr, (const int*)blocks, d_out, n, d_density);
    OTM_HIP(hipGetLastError());
    long k = 0;
    OTM_HIP(hipMemcpy(&k, total, sizeof(long), hipMemcpyDeviceToHost));
    *kept = k;
    return OTM_OK;
  } catch (const HipError& e) {
    return fail(e);
  }
}

}  // extern "C"
