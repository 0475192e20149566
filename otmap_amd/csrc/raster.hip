// rasterize_image, per-face density mode (otlib/utils/rasterizer.cpp:138-237,
// 305-308) on the GPU.
//
// Reference semantics: faces in index order; for each face the pixel box
// [int(min*R), int(max*R)+1) clamped to the image, scanned row by row; a
// pixel is covered when all four inverse-bilinear weights of its centre lie
// in [0, 1]; covered pixels are overwritten with mass/area, mass =
// float(1/F) * density(f), area = |signed area of the quad|.  The last face
// (largest index) covering a pixel therefore wins.
//
// Where the bilinear solve fails (D < 0) the reference tests an uninitialised
// vector that, in its Release build, still holds the weights of the last
// pixel whose solve succeeded -- in the same face or, for the first pixels of
// a face, in the preceding faces.  The passes reproduce that: pass 0 records
// per face the weights of its last successful solve in scan order (its "exit
// weights", if any); a thread of the cover pass that meets a failing pixel
// before any success in its own face takes the exit weights of the nearest
// preceding face that has them, up to MAX_WALK faces back (beyond that the
// pixel counts as not covered; this never triggers on meshes whose faces
// cover pixels).
//
// Pass 0 (exit): one thread per face.  Pass 1 (cover): 16x16 faces per
// workgroup, one per thread; the winners of the workgroup's pixel footprint
// are accumulated with atomicMax in an LDS tile and merged into the global
// winner image (global atomics directly when the footprint exceeds the
// tile).  Pass 2 (resolve): one thread per pixel.

#include <cfloat>
#include <mutex>

#include "../../include/otmap_capi.h"
#include "geometry.h"
#include "otm_internal.h"

using namespace otm;
using namespace otm::geo;

namespace otm {
namespace {

constexpr int FB = 16;              // faces per workgroup side
constexpr int LDS_PIX = 64 * 64;    // winner tile capacity (pixels)
constexpr int MAX_WALK = 64;
constexpr int BIG_FACE = 128;       // faces with larger pixel boxes go to k_raster_big

__device__ inline void quad_pts(const double* xy, int n, int f, P2* p) {
  const long V = long(n + 1) * (n + 1);
  const int i = f / n, j = f - (f / n) * n;
  const int v0 = j + i * (n + 1), v1 = j + (i + 1) * (n + 1), v2 = (j + 1) + (i + 1) * (n + 1),
            v3 = (j + 1) + i * (n + 1);
  p[0] = {xy[v0], xy[V + v0]};
  p[1] = {xy[v1], xy[V + v1]};
  p[2] = {xy[v2], xy[V + v2]};
  p[3] = {xy[v3], xy[V + v3]};
}

struct Box {
  int x0, x1, y0, y1;
};

// clamp of the reference's int casts (out-of-range values cannot index the image anyway)
__device__ inline int to_int(double v) {
  if (!(v > -2147483648.0)) return -2147483647;
  if (!(v < 2147483647.0)) return 2147483647;
  return int(v);
}

__device__ inline Box pixel_box(const P2* p, int rows, int cols) {
  double lox = p[0].x, loy = p[0].y, hix = p[0].x, hiy = p[0].y;
  for (int t = 1; t < 4; ++t) {
    lox = fmin(lox, p[t].x); loy = fmin(loy, p[t].y);
    hix = fmax(hix, p[t].x); hiy = fmax(hiy, p[t].y);
  }
  Box b;
  b.x0 = max(to_int(lox * double(cols)), 0);
  b.y0 = max(to_int(loy * double(rows)), 0);
  const int ex = to_int(hix * double(cols)), ey = to_int(hiy * double(rows));
  b.x1 = (ex >= cols - 1) ? cols : ex + 1;
  b.y1 = (ey >= rows - 1) ? rows : ey + 1;
  return b;
}

__device__ inline P2 pixel_centre(int x, int y, int rows, int cols) {
  return {(double(x) + 0.5) / double(cols), (double(y) + 0.5) / double(rows)};
}

// pass 0: exit weights of every face (the last successful solve of its scan,
// found by scanning its pixel box backwards); ok[f] = 0 when none succeeds
__global__ __launch_bounds__(256) void k_raster_exit(const double* __restrict__ xy, int n, int rows, int cols,
                                                    double* __restrict__ ew, int* __restrict__ ok) {
  const long f = long(blockIdx.x) * 256 + threadIdx.x;
  if (f >= long(n) * n) return;
  P2 p[4];
  quad_pts(xy, n, int(f), p);
  const Box b = pixel_box(p, rows, cols);
  double w[4];
  for (int y = b.y1 - 1; y >= b.y0; --y)
    for (int x = b.x1 - 1; x >= b.x0; --x)
      if (bilinear_weights(pixel_centre(x, y, rows, cols), p, w)) {
        for (int q = 0; q < 4; ++q) ew[4 * f + q] = w[q];
        ok[f] = 1;
        return;
      }
  ok[f] = 0;
}

// weights in effect after the scan of faces <= g (last successful solve), or false
__device__ bool exit_weights(const double* ew, const int* ok, int g, double* w) {
  for (int k = 0; k < MAX_WALK && g >= 0; ++k, --g)
    if (ok[g]) {
      for (int q = 0; q < 4; ++q) w[q] = ew[4 * long(g) + q];
      return true;
    }
  return false;
}

__device__ inline void atomic_max_int(int* p, int v) { atomicMax(p, v); }

// One face's scan of the pixels [k0, k1) of its box (row-major over the box),
// entering with the weights state (w, have) of the reference's loop.
__device__ inline void scan_pixels(const P2* p, int f, const Box& b, long k0, long k1, int rows, int cols,
                                   const double* ew, const int* ok, double* w, bool have, int* tile, int ux0,
                                   int uy0, int tw, int* winner) {
  const int bw = b.x1 - b.x0;
  for (long k = k0; k < k1; ++k) {
    const int y = b.y0 + int(k / bw), x = b.x0 + int(k % bw);
    if (bilinear_weights(pixel_centre(x, y, rows, cols), p, w)) {
      have = true;
    } else if (!have) {
      if (!exit_weights(ew, ok, f - 1, w)) { w[0] = w[1] = w[2] = w[3] = -1.; }
      have = true;
    }
    if (weights_in_unit(w)) {
      if (tile) atomic_max_int(&tile[(y - uy0) * tw + (x - ux0)], f);
      else atomic_max_int(&winner[long(y) * cols + x], f);
    }
  }
}

// Faces whose pixel box exceeds BIG_FACE pixels, one workgroup each: the box
// is cut into 256 runs in scan order; each thread first finds the last
// successful solve of its run (scanning it backwards), then enters its run
// with the weights of the nearest earlier run that has one -- the state the
// one-thread scan would carry -- and scans it forwards.
__global__ __launch_bounds__(256) void k_raster_big(const double* __restrict__ xy, int n, int rows, int cols,
                                                   const double* __restrict__ ew, const int* __restrict__ ok,
                                                   const int* __restrict__ big, const int* __restrict__ nbig,
                                                   int* __restrict__ winner) {
  __shared__ double rw[256][4];
  __shared__ int rok[256];
  const int t = threadIdx.x;
  const int count = *nbig;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int f = big[item];
    P2 p[4];
    quad_pts(xy, n, f, p);
    const Box b = pixel_box(p, rows, cols);
    const long area = long(b.x1 - b.x0) * (b.y1 - b.y0);
    const long run = (area + 255) / 256;
    const long k0 = min(area, t * run), k1 = min(area, k0 + run);
    const int bw = b.x1 - b.x0;
    double w[4];
    int found = 0;
    for (long k = k1 - 1; k >= k0; --k) {
      const int y = b.y0 + int(k / bw), x = b.x0 + int(k % bw);
      if (bilinear_weights(pixel_centre(x, y, rows, cols), p, w)) { found = 1; break; }
    }
    rok[t] = found;
    for (int q = 0; q < 4; ++q) rw[t][q] = found ? w[q] : -1.;
    __syncthreads();
    double win[4] = {-1., -1., -1., -1.};
    bool have = false;
    for (int s = t - 1; s >= 0; --s)
      if (rok[s]) {
        for (int q = 0; q < 4; ++q) win[q] = rw[s][q];
        have = true;
        break;
      }
    scan_pixels(p, f, b, k0, k1, rows, cols, ew, ok, win, have, nullptr, 0, 0, 0, winner);
    __syncthreads();  // rw / rok are rewritten by the next item
  }
}

__global__ __launch_bounds__(256) void k_raster_cover(const double* __restrict__ xy, int n, int rows, int cols,
                                                     const double* __restrict__ ew, const int* __restrict__ ok,
                                                     int* __restrict__ big, int* __restrict__ nbig,
                                                     int* __restrict__ winner) {
  __shared__ int tile[LDS_PIX];
  __shared__ int ux0, ux1, uy0, uy1;
  const int fi = blockIdx.y * FB + threadIdx.y, fj = blockIdx.x * FB + threadIdx.x;
  const int tid = threadIdx.y * FB + threadIdx.x;
  const bool valid = fi < n && fj < n;
  const int f = valid ? fj + fi * n : -1;
  P2 p[4];
  Box b = {0, 0, 0, 0};
  bool mine = false;  // scanned here (not deferred to k_raster_big)
  if (valid) {
    quad_pts(xy, n, f, p);
    b = pixel_box(p, rows, cols);
    const bool nonempty = b.x1 > b.x0 && b.y1 > b.y0;
    if (nonempty && long(b.x1 - b.x0) * (b.y1 - b.y0) > BIG_FACE) {
      big[atomicAdd(nbig, 1)] = f;
    } else {
      mine = true;
    }
  }
  if (tid == 0) { ux0 = 2147483647; uy0 = 2147483647; ux1 = -2147483647; uy1 = -2147483647; }
  __syncthreads();
  if (mine && b.x1 > b.x0 && b.y1 > b.y0) {
    atomicMin(&ux0, b.x0); atomicMin(&uy0, b.y0);
    atomicMax(&ux1, b.x1); atomicMax(&uy1, b.y1);
  }
  __syncthreads();
  const int tw = ux1 - ux0, th = uy1 - uy0;
  const bool use_lds = tw > 0 && th > 0 && long(tw) * th <= LDS_PIX;
  if (use_lds)
    for (int k = tid; k < tw * th; k += 256) tile[k] = -1;
  __syncthreads();
  if (mine) {
    double w[4] = {-1., -1., -1., -1.};
    const long area = long(max(b.x1 - b.x0, 0)) * max(b.y1 - b.y0, 0);
    scan_pixels(p, f, b, 0, area, rows, cols, ew, ok, w, false, use_lds ? tile : nullptr, ux0, uy0, tw, winner);
  }
  __syncthreads();
  if (use_lds)
    for (int k = tid; k < tw * th; k += 256) {
      const int v = tile[k];
      if (v >= 0) atomic_max_int(&winner[long(uy0 + k / tw) * cols + ux0 + k % tw], v);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_raster_resolve(const double* __restrict__ xy, const double* __restrict__ density,
                                                       int n, int rows, int cols, const int* __restrict__ winner,
                                                       T* __restrict__ img) {
  const long k = long(blockIdx.x) * 256 + threadIdx.x;
  if (k >= long(rows) * cols) return;
  const int f = winner[k];
  double v = 0.;
  if (f >= 0) {
    P2 p[4];
    quad_pts(xy, n, f, p);
    const double area = fabs(sarea(p[0], p[1], p[2]) + sarea(p[0], p[2], p[3]));
    const float scale = float(1. / double(long(n) * n));
    const double mass = double(scale) * (density ? density[f] : 1.0);
    v = mass / area;
  }
  img[k] = T(v);
}

__global__ void k_fill_int(int* p, long count, int v) {
  for (long i = long(blockIdx.x) * 256 + threadIdx.x; i < count; i += long(gridDim.x) * 256) p[i] = v;
}

}  // namespace

// Exit-weight records of the faces (4 doubles + a flag per face): one buffer
// per device, kept and grown by the library; calls that overlap on one
// device serialise on its lock, which is held until the passes complete.
struct RasterScratch {
  std::mutex mu;
  double* ew = nullptr;
  int* ok = nullptr;
  int* big = nullptr;  // deferred faces (F) + their count
  int* nbig = nullptr;
  long cap = 0;
};
RasterScratch& raster_scratch(int device) {
  static RasterScratch per_device[64];
  return per_device[device & 63];
}

// winner image (rows*cols ints) + image; img_f32 selects float output
void launch_rasterize(const double* d_xy, const double* d_density, int n, int rows, int cols, int* d_winner,
                      void* d_img, int img_f32, hipStream_t s) {
  int device = 0;
  OTM_HIP(hipGetDevice(&device));
  RasterScratch& rs = raster_scratch(device);
  std::lock_guard<std::mutex> lock(rs.mu);
  const long F = long(n) * n;
  if (F > rs.cap) {
    OTM_HIP(hipStreamSynchronize(s));
    if (rs.ew) (void)hipFree(rs.ew);
    if (rs.ok) (void)hipFree(rs.ok);
    if (rs.big) (void)hipFree(rs.big);
    rs.ew = nullptr;
    rs.ok = nullptr;
    rs.big = nullptr;
    rs.nbig = nullptr;
    rs.cap = 0;
    OTM_HIP(hipMalloc(&rs.ew, size_t(F) * 4 * sizeof(double)));
    OTM_HIP(hipMalloc(&rs.ok, size_t(F) * sizeof(int)));
    OTM_HIP(hipMalloc(&rs.big, size_t(F + 1) * sizeof(int)));
    rs.nbig = rs.big + F;
    rs.cap = F;
  }
  const long P = long(rows) * cols;
  const unsigned fill_blocks = unsigned(std::min<long>((P + 255) / 256, 8192));
  hipLaunchKernelGGL(k_fill_int, dim3(fill_blocks), dim3(256), 0, s, d_winner, P, -1);
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, s, rs.nbig, 1L, 0);
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_raster_exit, dim3(unsigned((F + 255) / 256)), dim3(256), 0, s, d_xy, n, rows, cols, rs.ew,
                     rs.ok);
  OTM_HIP(hipGetLastError());
  const dim3 grid((n + FB - 1) / FB, (n + FB - 1) / FB);
  hipLaunchKernelGGL(k_raster_cover, grid, dim3(FB, FB), 0, s, d_xy, n, rows, cols, (const double*)rs.ew,
                     (const int*)rs.ok, rs.big, rs.nbig, d_winner);
  OTM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_raster_big, dim3(unsigned(std::min<long>(F, 1024))), dim3(256), 0, s, d_xy, n, rows, cols,
                     (const double*)rs.ew, (const int*)rs.ok, (const int*)rs.big, (const int*)rs.nbig, d_winner);
  OTM_HIP(hipGetLastError());
  const unsigned rb = unsigned((P + 255) / 256);
  if (img_f32)
    hipLaunchKernelGGL(k_raster_resolve<float>, dim3(rb), dim3(256), 0, s, d_xy, d_density, n, rows, cols,
                       (const int*)d_winner, (float*)d_img);
  else
    hipLaunchKernelGGL(k_raster_resolve<double>, dim3(rb), dim3(256), 0, s, d_xy, d_density, n, rows, cols,
                       (const int*)d_winner, (double*)d_img);
  OTM_HIP(hipGetLastError());
  OTM_HIP(hipStreamSynchronize(s));  // the scratch records are free again
}

}  // namespace otm

extern "C" {

int otm_rasterize_device(int device, int n, const double* d_mesh_xy, const double* d_density, int rows, int cols,
                         int img_f32, void* d_img, int32_t* d_winner) {
  if (n < 1 || rows < 1 || cols < 1 || !d_mesh_xy || !d_img || !d_winner) {
    set_error("invalid rasterize arguments");
    return OTM_ERR_ARG;
  }
  try {
    OTM_HIP(hipSetDevice(device));
    launch_rasterize(d_mesh_xy, d_density, n, rows, cols, d_winner, d_img, img_f32, nullptr);
    OTM_HIP(hipStreamSynchronize(nullptr));
    return OTM_OK;
  } catch (const HipError& e) {
    set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
    return OTM_ERR_HIP;
  }
}

int otm_rasterize(int device, int n, const double* mesh_xy, const double* density, int rows, int cols, double* img) {
  if (n < 1 || rows < 1 || cols < 1 || !mesh_xy || !img) {
    set_error("invalid rasterize arguments");
    return OTM_ERR_ARG;
  }
  double *dxy = nullptr, *dd = nullptr, *dimg = nullptr;
  int* dw = nullptr;
  int rc = OTM_OK;
  try {
    OTM_HIP(hipSetDevice(device));
    const size_t V = size_t(n + 1) * (n + 1), F = size_t(n) * n, P = size_t(rows) * cols;
    OTM_HIP(hipMalloc(&dxy, 2 * V * sizeof(double)));
    OTM_HIP(hipMalloc(&dimg, P * sizeof(double)));
    OTM_HIP(hipMalloc(&dw, P * sizeof(int)));
    OTM_HIP(hipMemcpy(dxy, mesh_xy, 2 * V * sizeof(double), hipMemcpyHostToDevice));
    if (density) {
      OTM_HIP(hipMalloc(&dd, F * sizeof(double)));
      OTM_HIP(hipMemcpy(dd, density, F * sizeof(double), hipMemcpyHostToDevice));
    }
    launch_rasterize(dxy, dd, n, rows, cols, dw, dimg, 0, nullptr);
    OTM_HIP(hipMemcpy(img, dimg, P * sizeof(double), hipMemcpyDeviceToHost));
  } catch (const HipError& e) {
    set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
    rc = OTM_ERR_HIP;
  }
  if (dxy) (void)hipFree(dxy);
  if (dd) (void)hipFree(dd);
  if (dimg) (void)hipFree(dimg);
  if (dw) (void)hipFree(dw);
  return rc;
}

}  // extern "C"
