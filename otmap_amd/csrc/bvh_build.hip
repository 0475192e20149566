// BVH over the forward quads built on the device, node for node and face
// slot for face slot the tree of the reference's BVH2D::build
// (otlib/utils/bvh2d.cpp:29-58, 167-243): face centroids; per node the AABB of its faces' vertices
// inflated by 2 eps diag + eps; leaf when <= leaf_size faces or depth >=
// max_depth; otherwise the longest axis is split at the float-rounded box
// midpoint by the reference's in-place two-pointer partition, and a split
// that leaves one side empty makes a leaf (after the partition's swaps).
//
// The build is level-synchronous: every node of one depth is processed at
// once, one thread per face slot.
//   * Boxes: order-free min / max, as atomics on order-preserving 64-bit keys
//     (one atomic per workgroup chunk in one node, else per run of one
//     node's slots in a wave).
//   * Partition: with L = the faces left of the split value (c < sv) in
//     [start, end), the two-pointer loop swaps the k-th right face in
//     [start, start + L) (counted from the left) with the k-th left face in
//     [start + L, end) (counted from the right) and nothing else; the ranks
//     come from two exclusive scans over all slots.  Its returned middle is
//     start + L, plus one when the face then at start + L lies exactly on
//     the split value (the loop's final `<=` test).
//   * Numbering: the reference appends a node's two children when it splits,
//     in depth-first order (first child's subtree first), so the split node
//     of preorder rank k (among split nodes) gets children 1 + 2k and 2 + 2k.
//     Ranks follow from the split-node count of each first child's subtree
//     (bottom-up), then top-down; escape links (the next node once a subtree
//     is done: first child -> second child -> parent's escape) likewise.
// Result: the same BNode array and face order as a sequential build of the
// reference's algorithm (tested against the oracle's).
//
// Every level up to the depth limit is enqueued at once: the range of each
// level (first record, node count) is written on the device by the previous
// level, node kernels run over the level's largest possible node count and
// read the actual one, so the host waits once, for the final node count.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <map>
#include <mutex>
#include <vector>

#include "otm_internal.h"

namespace otm {
namespace {

constexpr int BT = 256;  // threads per workgroup

typedef unsigned long long u64;

// order-preserving key of a double (min / max of keys = min / max of values)
__device__ inline u64 dkey(double x) {
  const u64 b = __double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ inline double dkey_inv(u64 k) {
  const u64 b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return __longlong_as_double(b);
}

struct Tmp {  // temporary node records, indexed by build order (level by level)
  int* start;
  int* end;
  int* child;   // first child's record, -1 for a leaf
  int* level;
  u64* key;     // 4 per node: min x, min y, max x, max y
  double* box;  // 4 per node: inflated box
  int* dim;
  float* sv;
  int* split;   // 1: the node splits (set before the partition, cleared if a side stays empty)
  int* nleft;   // L
  int* mid;
  int* nsplit;  // split nodes in the subtree (the node included)
  int* rank;    // preorder rank among split nodes
  int* fin;     // final index
  int* esc;     // final escape index
};

__device__ inline void quad_v(int n, int f, int* v) {
  const int i = f / n, j = f - (f / n) * n;
  v[0] = j + i * (n + 1);
  v[1] = j + (i + 1) * (n + 1);
  v[2] = (j + 1) + (i + 1) * (n + 1);
  v[3] = (j + 1) + i * (n + 1);
}

// centroids in the reference's order: ((((0 + p0) + p1) + p2) + p3) / 4
__global__ __launch_bounds__(BT) void k_centroids(int n, const double* __restrict__ xy, int* __restrict__ fid,
                                                 double* __restrict__ cx, double* __restrict__ cy,
                                                 int* __restrict__ pn) {
  const long F = long(n) * n, V = long(n + 1) * (n + 1);
  for (long f = long(blockIdx.x) * BT + threadIdx.x; f < F; f += long(gridDim.x) * BT) {
    int v[4];
    quad_v(n, int(f), v);
    double x = 0., y = 0.;
    for (int q = 0; q < 4; ++q) {
      x += xy[v[q]];
      y += xy[V + v[q]];
    }
    cx[f] = x / 4;
    cy[f] = y / 4;
    fid[f] = int(f);
    pn[f] = 0;
  }
}

// Level table lv (device): lv[2 l] = first record of level l, lv[2 l + 1] =
// its node count, lv[LV_ERR] != 0 when a level would not fit the records.
// Every level of the depth limit is enqueued at once; node kernels read
// their level's range from the table (levels past the last one are empty).
__device__ inline int lv_err_index(int levels) { return 2 * levels; }

__global__ __launch_bounds__(BT) void k_root(Tmp T, int F, int* __restrict__ lv, int levels) {
  if (blockIdx.x != 0) return;
  for (int i = threadIdx.x; i <= 2 * levels; i += BT) lv[i] = 0;
  __syncthreads();
  if (threadIdx.x != 0) return;
  lv[1] = 1;  // level 0: the root alone
  T.start[0] = 0;
  T.end[0] = F;
  T.level[0] = 0;
}

__global__ __launch_bounds__(BT) void k_lvl_init(Tmp T, const int* __restrict__ lv, int l) {
  const int k = blockIdx.x * BT + threadIdx.x;
  if (k >= lv[2 * l + 1]) return;
  const int id = lv[2 * l] + k;
  T.key[4 * id + 0] = ~0ull;
  T.key[4 * id + 1] = ~0ull;
  T.key[4 * id + 2] = 0ull;
  T.key[4 * id + 3] = 0ull;
  T.nleft[id] = 0;
  T.child[id] = -1;
}

// A workgroup handles CHUNK consecutive face slots, PER per thread (slot
// c0 + t + k BT).  Node ranges are contiguous, so when the chunk's first and
// last slots belong to one node (not to finished leaves, which may enclose
// active nodes) every slot between does: the workgroup then
// reduces in registers and LDS and issues one atomic per value; otherwise
// (chunks across node boundaries) each wave reduces the runs of slots of one
// node by a segmented scan and issues one atomic per value per run.
constexpr int PER = 4;
constexpr int CHUNK = BT * PER;

template <class Op>
__device__ inline double block_reduce(double v, Op op, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < BT / 64; ++w) r = op(r, sh[w]);
  return r;
}

// inclusive segmented scan of op over the wave's lanes, segments = runs of
// equal node ids (runs are contiguous: the lanes hold consecutive slots);
// the last lane of a run ends with the run's total
template <class T, class Op>
__device__ inline T seg_scan(T v, int nd, Op op) {
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const T u = __shfl_up(v, o);
    const int m = __shfl_up(nd, o);
    if (lane >= o && m == nd) v = op(v, u);
  }
  return v;
}
__device__ inline bool run_end(int nd) {
  const int lane = threadIdx.x & 63;
  const int nx = __shfl_down(nd, 1);
  return lane == 63 || nx != nd;
}

__device__ inline void face_box(int n, long V, const double* __restrict__ xy, int f, double& lx, double& ly,
                                double& hx, double& hy) {
  int v[4];
  quad_v(n, f, v);
  for (int q = 0; q < 4; ++q) {
    const double x = xy[v[q]], y = xy[V + v[q]];
    lx = fmin(lx, x); ly = fmin(ly, y);
    hx = fmax(hx, x); hy = fmax(hy, y);
  }
}

// node boxes: min / max of the vertices of every face of the node
__global__ __launch_bounds__(BT) void k_lvl_box(int n, const double* __restrict__ xy, const int* __restrict__ fid,
                                               const int* __restrict__ pn, Tmp T) {
  __shared__ double sh[BT / 64];
  const long F = long(n) * n, V = long(n + 1) * (n + 1);
  const long c0 = long(blockIdx.x) * CHUNK, c1 = c0 + CHUNK < F ? c0 + CHUNK : F;
  const int n0 = pn[c0], n1 = pn[c1 - 1];
  if (n0 == n1 && n0 >= 0) {  // one node (finished slots at both ends prove nothing)
    double lx = DBL_MAX, ly = DBL_MAX, hx = -DBL_MAX, hy = -DBL_MAX;
    for (int k = 0; k < PER; ++k) {
      const long p = c0 + threadIdx.x + k * BT;
      if (p < c1) face_box(n, V, xy, fid[p], lx, ly, hx, hy);
    }
    auto mn = [](double a, double b) { return fmin(a, b); };
    auto mx = [](double a, double b) { return fmax(a, b); };
    lx = block_reduce(lx, mn, sh);
    ly = block_reduce(ly, mn, sh);
    hx = block_reduce(hx, mx, sh);
    hy = block_reduce(hy, mx, sh);
    if (threadIdx.x == 0) {
      atomicMin(&T.key[4 * n0 + 0], dkey(lx));
      atomicMin(&T.key[4 * n0 + 1], dkey(ly));
      atomicMax(&T.key[4 * n0 + 2], dkey(hx));
      atomicMax(&T.key[4 * n0 + 3], dkey(hy));
    }
    return;
  }
  auto mn = [](double a, double b) { return fmin(a, b); };
  auto mx = [](double a, double b) { return fmax(a, b); };
  for (int k = 0; k < PER; ++k) {
    const long p = c0 + threadIdx.x + k * BT;
    const int nd = p < c1 ? pn[p] : -1;
    double lx = DBL_MAX, ly = DBL_MAX, hx = -DBL_MAX, hy = -DBL_MAX;
    if (nd >= 0) face_box(n, V, xy, fid[p], lx, ly, hx, hy);
    if (__all(nd < 0)) continue;
    lx = seg_scan(lx, nd, mn);
    ly = seg_scan(ly, nd, mn);
    hx = seg_scan(hx, nd, mx);
    hy = seg_scan(hy, nd, mx);
    if (run_end(nd) && nd >= 0) {  // one atomic per value per run of the wave
      atomicMin(&T.key[4 * nd + 0], dkey(lx));
      atomicMin(&T.key[4 * nd + 1], dkey(ly));
      atomicMax(&T.key[4 * nd + 2], dkey(hx));
      atomicMax(&T.key[4 * nd + 3], dkey(hy));
    }
  }
}

// inflated box, leaf test, split axis and value (otlib/utils/bvh2d.cpp:193-221)
__global__ __launch_bounds__(BT) void k_lvl_split(Tmp T, const int* __restrict__ lv, int l, int leaf_size,
                                                 int max_depth) {
  const int k = blockIdx.x * BT + threadIdx.x;
  if (k >= lv[2 * l + 1]) return;
  const int id = lv[2 * l] + k;
  const double lo[2] = {dkey_inv(T.key[4 * id + 0]), dkey_inv(T.key[4 * id + 1])};
  const double hi[2] = {dkey_inv(T.key[4 * id + 2]), dkey_inv(T.key[4 * id + 3])};
  const double eps = DBL_EPSILON;
  const double diag[2] = {hi[0] - lo[0], hi[1] - lo[1]};
  double blo[2], bhi[2];
  for (int d = 0; d < 2; ++d) {
    const double pad = diag[d] * (2. * eps) + eps;
    blo[d] = lo[d] - pad;
    bhi[d] = hi[d] + pad;
  }
  T.box[4 * id + 0] = blo[0];
  T.box[4 * id + 1] = blo[1];
  T.box[4 * id + 2] = bhi[0];
  T.box[4 * id + 3] = bhi[1];
  const int cnt = T.end[id] - T.start[id];
  const bool leaf = cnt <= leaf_size || T.level[id] >= max_depth;
  const int dim = diag[1] > diag[0] ? 1 : 0;
  T.dim[id] = dim;
  T.sv[id] = float(0.5 * (bhi[dim] + blo[dim]));
  T.split[id] = leaf ? 0 : 1;
}

// left flags (centroid < split value) and L per splitting node (chunks as k_lvl_box)
__global__ __launch_bounds__(BT) void k_lvl_left(int F, const double* __restrict__ cx, const double* __restrict__ cy,
                                                const int* __restrict__ pn, Tmp T, unsigned char* __restrict__ left) {
  __shared__ double sh[BT / 64];
  const long c0 = long(blockIdx.x) * CHUNK, c1 = c0 + CHUNK < F ? c0 + CHUNK : F;
  const int n0 = pn[c0], n1 = pn[c1 - 1];
  const bool one = n0 == n1 && n0 >= 0;
  int tot = 0;
  for (int k = 0; k < PER; ++k) {
    const long p = c0 + threadIdx.x + k * BT;
    int nd = p < c1 ? pn[p] : -1;
    if (nd >= 0 && !T.split[nd]) nd = -1;
    int l = 0;
    if (nd >= 0) {
      const double c = T.dim[nd] ? cy[p] : cx[p];
      l = c < double(T.sv[nd]) ? 1 : 0;
    }
    if (p < c1) left[p] = (unsigned char)l;
    if (one) {
      tot += l;
      continue;
    }
    if (__all(nd < 0)) continue;
    const int s = seg_scan(l, nd, [](int a, int b) { return a + b; });
    if (run_end(nd) && nd >= 0 && s) atomicAdd(&T.nleft[nd], s);
  }
  if (one && T.split[n0]) {
    const double s = block_reduce(double(tot), [](double a, double b) { return a + b; }, sh);
    if (threadIdx.x == 0 && s > 0.) atomicAdd(&T.nleft[n0], int(s));
  }
}

// faces on the wrong side of start + L: right faces before it (bl), left faces from it on (br)
__global__ __launch_bounds__(BT) void k_lvl_bad(int F, const int* __restrict__ pn, Tmp T,
                                               const unsigned char* __restrict__ left, int* __restrict__ bl,
                                               int* __restrict__ br) {
  const long p = long(blockIdx.x) * BT + threadIdx.x;
  if (p > F) return;
  int a = 0, c = 0;
  if (p < F) {
    const int nd = pn[p];
    if (nd >= 0 && T.split[nd]) {
      const bool region = p < T.start[nd] + T.nleft[nd];
      a = (region && !left[p]) ? 1 : 0;
      c = (!region && left[p]) ? 1 : 0;
    }
  }
  bl[p] = a;  // slot F stays 0: scans give totals at F
  br[p] = c;
}

// slot of the k-th misplaced left face of each node, counted from the node's end
__global__ __launch_bounds__(BT) void k_lvl_rpos(int F, const int* __restrict__ pn, Tmp T,
                                                const int* __restrict__ br, const int* __restrict__ sr,
                                                int* __restrict__ rpos) {
  const long p = long(blockIdx.x) * BT + threadIdx.x;
  if (p >= F || !br[p]) return;
  const int nd = pn[p];
  const int s0 = sr[T.start[nd]], e0 = sr[T.end[nd]];
  const int r = e0 - sr[p] - 1;  // misplaced left faces after p in the node
  rpos[s0 + r] = int(p);
}

// the two-pointer swaps: the k-th misplaced right face (from the left) with
// the k-th misplaced left face (from the right)
__global__ __launch_bounds__(BT) void k_lvl_swap(int F, const int* __restrict__ pn, Tmp T, const int* __restrict__ bl,
                                                const int* __restrict__ sl, const int* __restrict__ sr,
                                                const int* __restrict__ rpos, int* __restrict__ fid,
                                                double* __restrict__ cx, double* __restrict__ cy) {
  const long p = long(blockIdx.x) * BT + threadIdx.x;
  if (p >= F || !bl[p]) return;
  const int nd = pn[p];
  const int r = sl[p] - sl[T.start[nd]];
  const long q = rpos[sr[T.start[nd]] + r];
  const int f = fid[p];
  fid[p] = fid[q];
  fid[q] = f;
  const double x = cx[p], y = cy[p];
  cx[p] = cx[q];
  cy[p] = cy[q];
  cx[q] = x;
  cy[q] = y;
}

// the partition's middle; a split that leaves a side empty makes a leaf;
// child counts for the next level (zero from K to the level's bound, so a
// scan over the bound + 1 entries gives the total at index K)
__global__ __launch_bounds__(BT) void k_lvl_mid(Tmp T, const int* __restrict__ lv, int l, int bound,
                                               const double* __restrict__ cx, const double* __restrict__ cy,
                                               int* __restrict__ nchild) {
  const int k = blockIdx.x * BT + threadIdx.x;
  if (k > bound) return;
  if (k >= lv[2 * l + 1]) {
    nchild[k] = 0;
    return;
  }
  const int id = lv[2 * l] + k;
  int two = 0;
  if (T.split[id]) {
    const int s = T.start[id], e = T.end[id];
    int mid = s + T.nleft[id];
    if (mid < e) {
      const double c = T.dim[id] ? cy[mid] : cx[mid];
      if (c == double(T.sv[id])) ++mid;
    }
    T.mid[id] = mid;
    if (mid == s || mid == e)
      T.split[id] = 0;
    else
      two = 2;
  }
  nchild[k] = two;
}

// child records of the next level (first children at b2 = b + K + exclusive
// scan); the next level's range goes to the table.  A binary tree over F
// faces with non-empty leaves has at most 2F - 1 records; a level that would
// pass `records` is not written and raises the table's error flag.
__global__ __launch_bounds__(BT) void k_lvl_children(Tmp T, int* __restrict__ lv, int l, int levels, long records,
                                                    const int* __restrict__ off) {
  const int k = blockIdx.x * BT + threadIdx.x;
  const int b = lv[2 * l], K = lv[2 * l + 1];
  const int b2 = b + K;
  const bool fits = long(b2) + off[K] <= records;
  if (k == 0 && l + 1 < levels) {
    lv[2 * (l + 1)] = b2;
    lv[2 * (l + 1) + 1] = fits ? off[K] : 0;
  }
  if (k == 0 && (!fits || (l + 1 == levels && off[K] > 0))) lv[lv_err_index(levels)] = 1;
  if (k >= K) return;
  const int id = b + k;
  if (!fits) {  // never for a well-formed tree: the level's nodes end as leaves
    T.split[id] = 0;
    return;
  }
  if (!T.split[id]) return;
  const int c = b2 + off[k];
  T.child[id] = c;
  T.start[c] = T.start[id];
  T.end[c] = T.mid[id];
  T.start[c + 1] = T.mid[id];
  T.end[c + 1] = T.end[id];
  T.level[c] = T.level[c + 1] = T.level[id] + 1;
}

// node of every slot at the next level (-1: its leaf is final)
__global__ __launch_bounds__(BT) void k_lvl_assign(int F, int* __restrict__ pn, Tmp T) {
  const long p = long(blockIdx.x) * BT + threadIdx.x;
  if (p >= F) return;
  const int nd = pn[p];
  if (nd < 0) return;
  if (!T.split[nd]) {
    pn[p] = -1;
    return;
  }
  pn[p] = T.child[nd] + (p < T.mid[nd] ? 0 : 1);
}

// split nodes per subtree, one level (bottom-up)
__global__ __launch_bounds__(BT) void k_count_split(Tmp T, const int* __restrict__ lv, int l) {
  const int k = blockIdx.x * BT + threadIdx.x;
  if (k >= lv[2 * l + 1]) return;
  const int id = lv[2 * l] + k;
  const int c = T.child[id];
  T.nsplit[id] = c < 0 ? 0 : 1 + T.nsplit[c] + T.nsplit[c + 1];
}

// preorder ranks, final indices and escape links, one level (top-down)
__global__ __launch_bounds__(BT) void k_number(Tmp T, const int* __restrict__ lv, int l) {
  const int k = blockIdx.x * BT + threadIdx.x;
  if (k >= lv[2 * l + 1]) return;
  const int id = lv[2 * l] + k;
  if (id == 0) {
    T.rank[0] = 0;
    T.fin[0] = 0;
    T.esc[0] = -1;
  }
  const int c = T.child[id];
  if (c < 0) return;
  const int r = T.rank[id];
  T.rank[c] = r + 1;
  T.rank[c + 1] = r + 1 + T.nsplit[c];
  T.fin[c] = 1 + 2 * r;
  T.fin[c + 1] = 2 + 2 * r;
  T.esc[c] = 1 + 2 * r + 1;
  T.esc[c + 1] = T.esc[id];
}

__global__ __launch_bounds__(BT) void k_write_nodes(Tmp T, int total, BNode* __restrict__ nodes) {
  const int id = blockIdx.x * BT + threadIdx.x;
  if (id >= total) return;
  BNode nd;
  nd.lox = T.box[4 * id + 0];
  nd.loy = T.box[4 * id + 1];
  nd.hix = T.box[4 * id + 2];
  nd.hiy = T.box[4 * id + 3];
  const int c = T.child[id];
  nd.first = c < 0 ? T.start[id] : T.fin[c];
  nd.count = c < 0 ? max(0, T.end[id] - T.start[id]) : -1;
  nd.escape = T.esc[id];
  nd.pad = 0;
  nodes[T.fin[id]] = nd;
}

inline unsigned blocks_for(long count) { return unsigned(std::max<long>(1, (count + BT - 1) / BT)); }

// Temporaries of the build, kept per device and grown to the largest grid
// built so far (a build on a 1024^2 grid needs ~0.3 GB, 2048^2 ~1.2 GB), so
// repeated builds allocate nothing.  Builds on one device are serialised.
struct Work {
  long F = 0;
  int levels = 0;
  size_t scan_bytes = 0;
  void* blob = nullptr;
  double *cx, *cy;
  int *pn, *bl, *br, *sl, *sr, *rpos, *nchild, *noff, *lv;
  unsigned char *left, *scan_tmp;
  Tmp T;
};

std::mutex g_work_mu;
std::map<int, Work> g_work;

Work& work_for(int device, long F, int levels) {
  Work& w = g_work[device];
  if (w.blob && w.F >= F && w.levels >= levels) return w;
  if (w.blob) (void)hipFree(w.blob);
  w = Work();
  const long NMAX = 2 * F + 2;
  size_t scan_bytes = 0;
  OTM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int*)nullptr, (int*)nullptr, int(F + 1)));
  // (byte offset, pointer) pairs filled in two passes: sizes, then addresses
  std::vector<std::pair<void**, size_t>> pieces;
  Tmp& T = w.T;
  auto add = [&](auto*& p, size_t count) { pieces.push_back({reinterpret_cast<void**>(&p), count * sizeof(*p)}); };
  add(w.cx, F); add(w.cy, F); add(w.pn, F); add(w.bl, F + 1); add(w.br, F + 1); add(w.sl, F + 1); add(w.sr, F + 1);
  add(w.rpos, F); add(w.left, F); add(w.nchild, F + 1); add(w.noff, F + 2); add(w.lv, 2 * levels + 2);
  add(w.scan_tmp, scan_bytes);
  add(T.start, NMAX); add(T.end, NMAX); add(T.child, NMAX); add(T.level, NMAX); add(T.key, 4 * NMAX);
  add(T.box, 4 * NMAX); add(T.dim, NMAX); add(T.sv, NMAX); add(T.split, NMAX); add(T.nleft, NMAX); add(T.mid, NMAX);
  add(T.nsplit, NMAX); add(T.rank, NMAX); add(T.fin, NMAX); add(T.esc, NMAX);
  size_t total = 0;
  std::vector<size_t> offs;
  for (const auto& pc : pieces) {
    total = (total + 255) & ~size_t(255);
    offs.push_back(total);
    total += std::max<size_t>(pc.second, 1);
  }
  OTM_HIP(hipMalloc(&w.blob, total));
  for (size_t i = 0; i < pieces.size(); ++i) *pieces[i].first = static_cast<char*>(w.blob) + offs[i];
  w.F = F;
  w.levels = levels;
  w.scan_bytes = scan_bytes;
  return w;
}

}  // namespace

// The whole build is enqueued without waiting on the device: every level up
// to the depth limit is launched (node kernels over the level's largest
// possible node count, min(2^l, F)), the level ranges stay in the device
// table, and the host reads the node count once at the end to size the
// node array.
void bvh_build_device(int n, const double* d_xy, int leaf_size, int max_depth, hipStream_t s, BNode** nodes_out,
                      int** faces_out, int* num_nodes) {
  const long F = long(n) * n;
  const long NMAX = 2 * F + 2;  // tree records: at most 2F - 1
  const int levels = max_depth + 1;  // a node at depth max_depth is a leaf
  int device = 0;
  OTM_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(g_work_mu);
  Work& w = work_for(device, F, levels);
  const Tmp& T = w.T;
  int* fid = nullptr;
  OTM_HIP(hipMalloc(&fid, size_t(F) * sizeof(int)));
  size_t scan_bytes = w.scan_bytes;

  const unsigned gF = blocks_for(F), gF1 = blocks_for(F + 1), gC = unsigned((F + CHUNK - 1) / CHUNK);
  hipLaunchKernelGGL(k_centroids, dim3(std::min<unsigned>(gF, 8192)), dim3(BT), 0, s, n, d_xy, fid, w.cx, w.cy,
                     w.pn);
  hipLaunchKernelGGL(k_root, dim3(1), dim3(BT), 0, s, T, int(F), w.lv, levels);
  std::vector<int> bound(levels);
  for (int l = 0; l < levels; ++l) {
    const int K = bound[l] = int(std::min<long>(F, l < 31 ? (1l << l) : F));
    const unsigned gK = blocks_for(K);
    hipLaunchKernelGGL(k_lvl_init, dim3(gK), dim3(BT), 0, s, T, (const int*)w.lv, l);
    hipLaunchKernelGGL(k_lvl_box, dim3(gC), dim3(BT), 0, s, n, d_xy, (const int*)fid, (const int*)w.pn, T);
    hipLaunchKernelGGL(k_lvl_split, dim3(gK), dim3(BT), 0, s, T, (const int*)w.lv, l, leaf_size, max_depth);
    hipLaunchKernelGGL(k_lvl_left, dim3(gC), dim3(BT), 0, s, int(F), (const double*)w.cx, (const double*)w.cy,
                       (const int*)w.pn, T, w.left);
    hipLaunchKernelGGL(k_lvl_bad, dim3(gF1), dim3(BT), 0, s, int(F), (const int*)w.pn, T,
                       (const unsigned char*)w.left, w.bl, w.br);
    OTM_HIP(hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, scan_bytes, w.bl, w.sl, int(F + 1), s));
    OTM_HIP(hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, scan_bytes, w.br, w.sr, int(F + 1), s));
    hipLaunchKernelGGL(k_lvl_rpos, dim3(gF), dim3(BT), 0, s, int(F), (const int*)w.pn, T, (const int*)w.br,
                       (const int*)w.sr, w.rpos);
    hipLaunchKernelGGL(k_lvl_swap, dim3(gF), dim3(BT), 0, s, int(F), (const int*)w.pn, T, (const int*)w.bl,
                       (const int*)w.sl, (const int*)w.sr, (const int*)w.rpos, fid, w.cx, w.cy);
    // child counts over the bound (zero past K), their scan, the next level
    hipLaunchKernelGGL(k_lvl_mid, dim3(blocks_for(long(K) + 1)), dim3(BT), 0, s, T, (const int*)w.lv, l, K,
                       (const double*)w.cx, (const double*)w.cy, w.nchild);
    size_t need = 0;
    OTM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, w.nchild, w.noff, K + 1, s));
    if (need > scan_bytes) throw HipError{hipErrorOutOfMemory, "bvh_build_device scan scratch"};
    OTM_HIP(hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, scan_bytes, w.nchild, w.noff, K + 1, s));
    hipLaunchKernelGGL(k_lvl_children, dim3(gK), dim3(BT), 0, s, T, w.lv, l, levels, NMAX, (const int*)w.noff);
    hipLaunchKernelGGL(k_lvl_assign, dim3(gF), dim3(BT), 0, s, int(F), w.pn, T);
    OTM_HIP(hipGetLastError());
  }
  for (int l = levels - 1; l >= 0; --l)
    hipLaunchKernelGGL(k_count_split, dim3(blocks_for(bound[l])), dim3(BT), 0, s, T, (const int*)w.lv, l);
  for (int l = 0; l < levels; ++l)
    hipLaunchKernelGGL(k_number, dim3(blocks_for(bound[l])), dim3(BT), 0, s, T, (const int*)w.lv, l);
  OTM_HIP(hipGetLastError());
  // the one wait of the build: node count (1 + 2 x split nodes) and the
  // level table's error flag
  int h[2] = {0, 0};
  OTM_HIP(hipMemcpyAsync(&h[0], T.nsplit, sizeof(int), hipMemcpyDeviceToHost, s));
  OTM_HIP(hipMemcpyAsync(&h[1], w.lv + 2 * levels, sizeof(int), hipMemcpyDeviceToHost, s));
  OTM_HIP(hipStreamSynchronize(s));
  if (h[1]) {
    (void)hipFree(fid);
    throw HipError{hipErrorInvalidValue, "bvh_build_device: tree larger than 2F records"};
  }
  const int total = 1 + 2 * h[0];
  BNode* nodes = nullptr;
  OTM_HIP(hipMalloc(&nodes, size_t(total) * sizeof(BNode)));
  hipLaunchKernelGGL(k_write_nodes, dim3(blocks_for(total)), dim3(BT), 0, s, T, total, nodes);
  OTM_HIP(hipGetLastError());
  OTM_HIP(hipStreamSynchronize(s));
  *nodes_out = nodes;
  *faces_out = fid;
  *num_nodes = total;
}

}  // namespace otm
