// Internal declarations shared by the HIP translation units of libotmap_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/otmap_capi.h"

namespace otm {

// ---------------------------------------------------------------------------
// error handling: every C entry point returns 0 or a negative code and
// records a message retrievable through otm_last_error().
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);

struct HipError {
  hipError_t code;
  std::string where;
};

#define OTM_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) throw ::otm::HipError{e_, std::string(#call) + " @ " __FILE__}; \
  } while (0)

// ---------------------------------------------------------------------------
// Face tiles used by every stencil kernel: TI rows (x index i) by TJ faces
// along the contiguous y index j; one workgroup of NT threads per tile.
// ---------------------------------------------------------------------------
constexpr int TI = 8;
constexpr int TJ = 64;
constexpr int NT = 256;

inline int tiles_j(int n) { return (n + TJ - 1) / TJ; }
inline int tiles_i(int n) { return (n + TI - 1) / TI; }
inline int num_tiles(int n) { return tiles_j(n) * tiles_i(n); }

// Reduction sites: each owns a region of the partials buffer so that a scalar
// kernel can fold partials written by two different producers.
enum Site { SITE_ADJUST = 0, SITE_RESID, SITE_DNORM, SITE_BETA, SITE_COEF, SITE_AUX, SITE_COUNT };
constexpr int SITE_VALUES = 8;  // values per workgroup a site can hold

// ---------------------------------------------------------------------------
// Device-resident solver state.  Written only by single-workgroup scalar
// kernels, read (gates, scalars) by the kernels that follow in stream order.
// ---------------------------------------------------------------------------
struct SolveState {
  double e;          // current residual |r|^2 / A (reference "residual")
  double alpha;      // last step length
  double beta;       // last conjugate-Jacobian factor
  double sum_r;      // sum of r_k (rank-1 pin correction of the Laplacian solve)
  double z[5];       // quartic coefficients of the last line search
  double dnorm;      // |d_hat| of the current iteration
  double threshold;
  double e0;         // initial residual
  double dens_scale; // adjust_density: 1 / integral
  double dens_shift; // adjust_density: beta' of the ratio cap
  double max_ratio;
  int it;            // completed iterations
  int max_iter;
  int cj;            // 1: conjugate-Jacobian beta, 0: beta = 0
  int done;          // 1: no further iteration starts
  int update;        // 1: the line-search update of the current iteration is pending
  int use_prev;      // 1: d = d_hat + beta * d_prev, 0: d = d_hat
  int exit_reason;   // 0 running, 1 threshold / max_iter, 2 |d_hat| <= 2*DBL_MIN
  int pad;
};

// Pointers read from device memory (job tables) are generic to the compiler;
// re-qualifying them as global lets it emit global instead of flat memory
// instructions (kernel pointer arguments are global already).
#define OTM_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline OTM_GLOBAL T* gp(T* p) {
  return (OTM_GLOBAL T*)p;
}

// Per-grid device buffers.
struct Fields {
  int n = 0;
  long F = 0, V = 0;
  double* rho = nullptr;     // adjusted density (F)
  double* psi = nullptr;     // potential (F); holds the potential between solves
  double* psi2 = nullptr;    // second buffer: iteration k reads psi_k from psi_buf(k), the update writes
                             // psi_{k+1} to psi_buf(k + 1) (neighbouring tiles read a tile's faces as their halo)
  double* g0 = nullptr;      // 2V (SoA): input staging of otm_line_search_coefficients (a solve forms grad psi from psi)
  double* r = nullptr;       // residual r_k (F), updated in place
  double* jd = nullptr;      // r_k - r_{k-1} (F)
  double* dhat = nullptr;    // Laplacian-solve direction (F)
  double* d[2] = {nullptr, nullptr};  // search direction ping-pong (F each)
  double** dpair = nullptr;  // device copy of {d[0], d[1]}
  double* a = nullptr;       // quartic coefficient a (F)
  double* b = nullptr;       // quartic coefficient b (F)
  double* partials = nullptr;  // SITE_COUNT regions of SITE_VALUES * pslots workgroup partials
  long pslots = 0;           // workgroups a site can hold
  double* tmp0 = nullptr;    // DCT intermediates (F each)
  double* tmp1 = nullptr;
  SolveState* state = nullptr;
  double* trace = nullptr;   // 3 * trace_cap: residual, alpha, beta per iteration
  int trace_cap = 0;
  // mixed-precision solve (allocated on first use): BF16 planes of the folded
  // right-hand side (2 x 3 x n*h), two stage intermediates (3 x n*n each) and
  // the FP64 unrefined solution
  unsigned short* mrhs = nullptr;
  unsigned short* mwa = nullptr;
  unsigned short* mwb = nullptr;
  double* x0 = nullptr;
  __host__ __device__ double* site(int s) const { return partials + long(s) * SITE_VALUES * pslots; }
  __host__ __device__ double* psi_buf(int k) const { return (k & 1) ? psi2 : psi; }
};

// ---------------------------------------------------------------------------
// Job tables.  Every kernel of the solve runs over a set of independent grids
// ("jobs") in one launch: workgroup b of a launch of kind K belongs to job p
// with starts[K][p] <= b < starts[K][p+1] and works on local tile
// b - starts[K][p] of that grid.  A single solve is a table of one job; a
// batch of uneven grids is one table, so each step of an iteration is one
// launch for the whole batch.
// ---------------------------------------------------------------------------
struct Job {
  Fields f;
  const double* P = nullptr;   // folded DCT plan of f.n (see DctPlan)
  const double* Pt = nullptr;
  const double* sq = nullptr;
  const unsigned short* Pp = nullptr;   // BF16 planes of P (3 x n*h) for the mixed solve
  const unsigned short* Ptp = nullptr;  // BF16 planes of Pt (3 x h*n)
  const double* ck = nullptr;           // cos(pi perm(r) / n), r < n, perm = even/odd order (mixed solve)
  const double* fft = nullptr;          // FFT route tables of f.n: twiddles e^{-2 pi i t/n}, then DCT phases e^{-i pi k/2n} (4n doubles)
  const double* dens_in = nullptr;  // input density of the current solve (F)
  double* fwd_out = nullptr;        // forward vertex positions (2V, SoA)
  double* adj_out = nullptr;        // copy of the adjusted density (F, nullable)
};

// Tile shapes (FM, FN, KB, G) of the folded FP64 DCT stages, picked per stage
// by the autotuner among those that divide n/2 for every job of the table:
// FM x FN outputs per workgroup, K steps of KB, G groups of four waves (G = 2:
// the two halves of K run on two groups, which needs n % 128 == 0 so that
// F1/F2 split at h/2 on a K-step boundary).
constexpr int NFOLD = 12;
constexpr int FOLD_CFG[NFOLD][4] = {{64, 64, 16, 1}, {64, 32, 16, 1}, {32, 64, 16, 1}, {32, 32, 16, 1},
                                    {64, 64, 32, 1}, {64, 32, 32, 1}, {32, 64, 32, 1}, {32, 32, 32, 1},
                                    {64, 64, 16, 2}, {64, 32, 16, 2}, {64, 64, 32, 2}, {64, 32, 32, 2}};
inline bool fold_cfg_valid(int n, int c) {
  const int h = n / 2;
  return n % 64 == 0 && h % FOLD_CFG[c][0] == 0 && h % FOLD_CFG[c][1] == 0 && h % FOLD_CFG[c][2] == 0 &&
         (FOLD_CFG[c][3] == 1 || n % 128 == 0);
}
// workgroups of folded stage st (0..3) with shape c for grid size n
inline int fold_tiles(int n, int st, int c) {
  const int h = n / 2;
  const int rows = st == 2 ? h : n, cols = st == 3 ? h : n;
  return (rows / FOLD_CFG[c][0]) * (cols / FOLD_CFG[c][1]);
}
// workgroups of the grid-stride ("flat") kernels per job
inline int flat_blocks(long F) { return int(std::min<long>((F + NT * 4 - 1) / (NT * 4), 1024)); }

// mixed-transform launch configurations: tile width (output columns per
// workgroup) and tile rows.  Configs 0, 1: the two-group form; configs 2, 3:
// the wide form (128 x 128 tiles, operands loaded straight into LDS) with
// four 64 x 64 waves or eight 64 x 32 waves.
constexpr int NMIXED = 4;
constexpr int MIXED_TN[NMIXED] = {32, 64, 128, 128};
constexpr int MIXED_TM[NMIXED] = {64, 64, 128, 128};

inline int mres_tiles(int n) { return (n / 32) * (n / 64); }
inline int upair_tiles(int n) { return tiles_i(n) * ((tiles_j(n) + 1) / 2); }
// Every job runs its Laplacian solve through one of three transforms, its
// route: the FP64 GEMM transform (route 0), the mixed BF16x3 transform with
// FP64 refinement (route 1) or the FP64 fast-DCT transform (route 2,
// fft.hip).  Launch kinds of one transform cover only the jobs routed to it.
// KIND_PLAIN64: one workgroup per 64 x 64 output tile (unfolded FP64 GEMMs);
// fold_kind(s, c) / mixed_kind(s, c): stage s of the folded FP64 / mixed
// transform with tile configuration c, over the jobs assigned configuration
// c for stage s (over every job for which c is valid in an autotuner table).
enum LaunchKind {
  KIND_STENCIL = 0,
  KIND_FLAT = 1,
  KIND_PLAIN64 = 2,
  KIND_MRES = 3,       // refinement right-hand side: one workgroup per 32 rows x 32 folded columns
  KIND_UPAIR = 4,      // update pass of the mixed form: one workgroup per stencil tile and its column mirror
  KIND_DNORM = 5,      // each job's workgroups in the last stage of its transform (writers of the |d_hat|^2 partials)
  KIND_STENCIL64 = 6,  // stencil tiles of the jobs on the FP64 transform (their plain update pass)
  KIND_FLATMX = 7,     // grid-stride blocks of the jobs on the mixed transform (first right-hand side)
  KIND_FFT = 8,        // fast-DCT passes of the FFT-route jobs with n <= FFT_SPLIT_N, packed: 1024/n sequences per workgroup
  KIND_FFT2 = 9,       // fast-DCT passes of the FFT-route jobs with n > FFT_SPLIT_N: one workgroup per 4 rows
  KIND_FFTW8 = 10,     // fast-DCT passes in row form, n = 1024: one wave per row, four per workgroup
  KIND_FFTW4 = 11,     // the same for n = 512
  KIND_FFTW16 = 12,    // the same for n = 2048
  KIND_M0 = 13,
  KIND_GEMM0 = KIND_M0 + 4 * NMIXED,
  KIND_COUNT = KIND_GEMM0 + 4 * NFOLD
};
inline int fold_kind(int stage, int c) { return KIND_GEMM0 + stage * NFOLD + c; }
inline int mixed_kind(int stage, int c) { return KIND_M0 + stage * NMIXED + c; }

// Tile configurations are assigned per job and stage (jcfg, jmcfg): jobs of
// different grid sizes in one batch each run the configuration tuned for
// their size, one launch per configuration in use.  Every configuration sums
// in the same order, so the assignment never changes results.
constexpr int FOLD_DEFAULT = 7;  // valid for every folded n
constexpr int MIXED_DEFAULT = 0;  // valid for every mixed n
struct JobSet {
  int count = 0;
  Job* jobs = nullptr;     // device copy of the table
  int* starts = nullptr;   // device, per kind: count + 1 workgroup prefix sums, then the job of each workgroup
  std::vector<Job> host;   // host copy
  std::vector<int> total;  // workgroups per kind
  bool fold = false;       // every job uses the folded transform
  bool tuning = false;     // autotuner table: every valid configuration covers every job
  std::vector<std::array<int, 4>> jcfg;   // per job: folded FP64 tile shape per stage
  std::vector<std::array<int, 4>> jmcfg;  // per job: mixed tile configuration per stage
  std::vector<int> jroute;                // per job: 0 FP64 GEMM transform, 1 mixed transform, 2 FP64 fast DCT
  int routed_mixed = 0;                   // jobs on the mixed transform
  int routed_fft = 0;                     // jobs on the fast-DCT transform
  int cfg[4] = {FOLD_DEFAULT, FOLD_DEFAULT, FOLD_DEFAULT, FOLD_DEFAULT};  // of the largest grid (reported)
  int mcfg[4] = {MIXED_DEFAULT, MIXED_DEFAULT, MIXED_DEFAULT, MIXED_DEFAULT};
  long version = 0;        // unique per table and assignment (captured graphs depend on it)
  std::vector<long> kofs;  // offset of each kind's block in `starts`
  const int* starts_of(int kind) const { return starts + kofs[kind]; }
};
// tuning = false: every job starts with the default configurations
void jobset_create(JobSet& js, const std::vector<Job>& jobs, bool tuning = false);
void jobset_assign(JobSet& js);  // after changing jcfg / jmcfg: new prefix sums, uploaded (synchronous)
// sets the transform route of every job (reassigns the table when it changes)
void jobset_route(JobSet& js, const std::vector<int>& route);
// Route of a grid of size n in the mixed precision mode: the BF16x3
// transform from mixed_min_n up (and only where it is supported), the FP64
// transform below, where it is both faster and exact.
inline int mixed_route(int n, int mixed_min_n) { return n % 64 == 0 && n >= 64 && n >= mixed_min_n ? 1 : 0; }
// Fast-DCT transform (fft.hip): n = 2^a 3^b, 8 <= n <= FFT_MAX_N, n % 4 == 0.
constexpr int FFT_MAX_N = 2048;
// Grid rows per workgroup of a fast-DCT pass (two per complex sequence): 2
// (256 threads, 48 n bytes of LDS, two or more workgroups per CU) up to
// FFT_SPLIT_N, 4 (512 threads, 80 n bytes) above, where two 2-row
// workgroups would no longer fit one CU's LDS.
constexpr int FFT_SPLIT_N = 1024;
constexpr int fft_rows(int n) { return n <= FFT_SPLIT_N ? 2 : 4; }
// The row form of the fast-DCT passes (fft.hip): one wave per grid row,
// whose DCT runs through an n/2-point complex FFT held in registers (n/128
// points per lane), four rows per workgroup (their transposed outputs leave
// as 32-byte runs).
constexpr bool fft_row_form(int n) { return n == 512 || n == 1024 || n == 2048; }
constexpr int FFT_ROW_W = 4;
// The packed workgroup form of the other sizes up to FFT_SPLIT_N (fft.hip):
// 256 threads, n/4 per complex sequence (two rows), so 1024 / n sequences
// per workgroup; its workgroups per grid.
constexpr bool fft_pack(int n) { return n <= FFT_SPLIT_N && !fft_row_form(n); }
constexpr int fft_pack_seqs(int n) { return FFT_SPLIT_N / n; }
constexpr int fft_pack_wgs(int n) { return (n / 2 + fft_pack_seqs(n) - 1) / fft_pack_seqs(n); }
// workgroups per grid of the fast-DCT passes (one |d_hat|^2 partial each)
constexpr int fft_wgs(int n) {
  return fft_row_form(n) ? n : fft_pack(n) ? fft_pack_wgs(n) : n / 4;
}
bool fft_supported(int n);
// FP64 route of a grid: the fast DCT when asked for and supported, else the GEMMs
inline int fp64_route(int n, int transform) { return transform == 1 && fft_supported(n) ? 2 : 0; }
void jobset_upload(JobSet& js, hipStream_t s);  // after changing js.host (same grid sizes)
void jobset_destroy(JobSet& js);

// Stencil kernel launchers (stencil.hip), one launch per step for all jobs
// Iteration driver constants (driver.hip): iterations per long graph replay,
// and the residual factor over the threshold below which a job is near its
// stop (single-iteration replays from then on).
constexpr int ITERS_PER_GRAPH = 4;
// near the stop: the residual of the last iterations falls by 5-10% per
// iteration on the benchmark inputs, so 2x the threshold is ~10 iterations out
constexpr double NEAR_FACTOR = 2.0;
// flags (one-job tables; null otherwise): the device address of the host's
// {open, far} pair, which the scalar kernels update when either changes
void launch_start(const JobSet& js, double max_ratio, int max_iter, double threshold, int cj, hipStream_t s,
                  int* flags = nullptr);
// k_beta + scalar kernel; the |d_hat|^2 partials of every job were written
// by the workgroups of the last stage of its transform (KIND_DNORM)
void launch_beta(const JobSet& js, hipStream_t s, int* flags = nullptr);
void launch_dir_coef(const JobSet& js, hipStream_t s, int* flags = nullptr);   // k_dir_coef + scalar kernel (line search)
void launch_dir_coef_pass(const JobSet& js, hipStream_t s);                     // k_dir_coef alone (kernel timing)
// jobs on the mixed transform run the folding update pass, which also writes
// the folded BF16 planes of their next right-hand side
void launch_update(const JobSet& js, hipStream_t s);
void launch_finish(const JobSet& js, hipStream_t s);     // forward positions, adjusted density copy
// d_open[0] = jobs whose state is not done, d_open[1] = those of them with a
// residual above NEAR_FACTOR x threshold and more than ITERS_PER_GRAPH
// iterations left before max_iter
void launch_count_open(const JobSet& js, int* d_open, hipStream_t s);
void launch_fill(double* p, long count, double v, hipStream_t s);
// component kernels (single evaluations)
void launch_vertex_gradients(const double* psi, double* G, int n, hipStream_t s);
void launch_residual_only(const double* psi, const double* rho, double* r, double* partials, int n, hipStream_t s);
void launch_params_1d(const double* g0, const double* dir, double* a, double* b, int n, hipStream_t s);
void launch_pick_steps(const double* z, int count, double* alpha, double* emin, hipStream_t s);

// DCT Laplacian solve (dct.hip): d_hat = M^+ (r - S e0), S = state->sum_r
struct DctPlan {
  int n = 0;
  double* C = nullptr;    // orthonormal DCT-II matrix, row-major C[k][i]
  double* Ct = nullptr;   // its transpose
  double* sq = nullptr;   // sin^2(pi m / (2n)), m = 0..2n
  // even/odd folded form (n % 64 == 0): with h = n/2,
  //   P[r][i] = C[2r][i] (r < h), C[2(r-h)+1][i] (r >= h), i < h   (n x h)
  //   Pt = P^T                                                      (h x n)
  bool fold = false;
  double* P = nullptr;
  double* Pt = nullptr;
  double* fft = nullptr;  // fast-DCT tables (fft.hip), when fft_supported(n)
};
void fft_plan_create(double*& table, int n);  // the 4n-double table of Job::fft, or nullptr when unsupported
// d_hat of every job on the fast-DCT route: three passes, fft_wgs(n) workgroups per job
void launch_fft_solve(const JobSet& js, bool gate, hipStream_t s);
// one of the three passes (0 rows forward, 1 columns, 2 rows inverse)
void launch_fft_pass(const JobSet& js, int pass, bool gate, hipStream_t s);
void dct_plan_create(DctPlan& p, int n);
void dct_plan_destroy(DctPlan& p);
// FP64 Laplacian solve of every job (folded when js.fold, else the unfolded
// transform of the single job with plan p)
void launch_dct_solve(const JobSet& js, const DctPlan* p, hipStream_t s);
void launch_dct_gemm_stage(const JobSet& js, const DctPlan* p, hipStream_t s);
// times every valid candidate tile shape of each folded stage over the jobs
// of each grid size of the table on stream s (cached per device, grid size
// and number of jobs of that size) and assigns the fastest to those jobs
void dct_tune(JobSet& js, int device, hipStream_t s);

// Picks the fastest of `count` launch configurations on stream s: after two
// warm-up launches of each valid configuration, TUNE_TRIALS rounds time
// TUNE_REPS back-to-back launches of every configuration in turn (so drift
// affects all alike); a configuration's time is its best round.  Ties go to
// the lower index; -1 when none is valid.
constexpr int TUNE_TRIALS = 5, TUNE_REPS = 6;
template <class Valid, class Launch>
int tune_pick(int count, Valid valid, Launch launch, hipStream_t s) {
  hipEvent_t e0, e1;
  OTM_HIP(hipEventCreate(&e0));
  OTM_HIP(hipEventCreate(&e1));
  std::vector<float> best(count, 1e30f);
  for (int c = 0; c < count; ++c)
    if (valid(c))
      for (int r = 0; r < 2; ++r) launch(c);
  for (int trial = 0; trial < TUNE_TRIALS; ++trial)
    for (int c = 0; c < count; ++c) {
      if (!valid(c)) continue;
      OTM_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < TUNE_REPS; ++r) launch(c);
      OTM_HIP(hipEventRecord(e1, s));
      OTM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      OTM_HIP(hipEventElapsedTime(&ms, e0, e1));
      best[c] = std::min(best[c], ms);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int pick = -1;
  for (int c = 0; c < count; ++c)
    if (valid(c) && (pick < 0 || best[c] < best[pick])) pick = c;
  return pick;
}

// Mixed-precision Laplacian solve (dct_mixed.hip): folded transform with
// 3-term BF16 split GEMMs on MFMA + FP64 refinement.  Requires n % 64 == 0.
struct MixedPlan {
  int n = 0;
  unsigned short* Pp = nullptr;   // BF16 planes of P  (3 x n*h)
  unsigned short* Ptp = nullptr;  // BF16 planes of Pt (3 x h*n)
  double* ck = nullptr;           // cos(pi perm(r) / n), r < n (eigenvalue factors, even/odd order)
};
bool mixed_supported(int n);
int mixed_tiles(int stage, int n, int cfg);      // workgroups of mixed stage 0..3 with config cfg
bool mixed_cfg_valid(int stage, int n, int cfg);  // n supported and the tile width divides the stage
void mixed_plan_create(MixedPlan& mp, const DctPlan& p);
void mixed_plan_destroy(MixedPlan& mp);
void fields_enable_mixed(Fields& f);  // allocates the mixed buffers (driver.hip)
// d_hat = M^+ (r - S e0) of every job of the table (all with mixed buffers and plans)
void launch_mixed_solve(const JobSet& js, int refine, bool gate, hipStream_t s, bool prep = true);
void launch_mprep(const JobSet& js, bool gate, hipStream_t s);  // folded planes of r' (first iteration)
void launch_mixed_gemm_stage(const JobSet& js, hipStream_t s);
// times every valid tile configuration of each mixed stage per grid size
// (cached and assigned like dct_tune)
void mixed_tune(JobSet& js, int device, hipStream_t s);

// Per-grid buffers and the iteration driver (driver.hip).  The per-iteration
// trace (residual, alpha, beta) keeps the first TRACE_MAX iterations whatever
// max_iter is (a solve may be given INT_MAX as "unlimited"); later iterations
// run but are not traced.
constexpr int TRACE_MAX = 4096;
inline int trace_cap_for(int max_iter) { return std::min(std::max(max_iter, 1), TRACE_MAX); }
void fields_create(Fields& f, int n, int trace_cap);
void fields_destroy(Fields& f);
bool fields_grow_trace(Fields& f, int cap);  // true when the buffer was reallocated

struct Driver {
  hipStream_t stream = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};     // captured iterations: a long and a one-iteration replay
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  long key = -1;                  // precision mode and tile shapes of the captured graphs
  // open / far-from-stop job counts, mapped pinned host memory: two slots of
  // two written by the count kernel after each replay of a batch; for a
  // one-job table the job's own {open, far} flags in slot 0, written by its
  // scalar kernels
  int* h_open = nullptr;
  int* d_open = nullptr;          // the device address of h_open
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evc[2] = {nullptr, nullptr};  // the copy of each slot is complete
};
void driver_create(Driver& d);
void driver_destroy(Driver& d);
void driver_invalidate(Driver& d);  // drop the captured graph (tables or shapes changed)
// one solve of every job of the table (js.host[p].dens_in / fwd_out / adj_out
// set by the caller; for mixed precision every job has its mixed buffers and
// plan); returns an OTM_* code, *ms = device time
int driver_solve(Driver& d, JobSet& js, const DctPlan* plan, const otm_options& opt, float* ms);

// Flattened BVH node of the inverse map (inverse.hip, bvh_build.hip):
// 48-byte records in the reference's node order.
struct alignas(16) BNode {
  double lox, loy, hix, hiy;
  int first;   // inner: first child; leaf: first face slot
  int count;   // leaf: number of faces; inner: -1
  int escape;  // next node after this subtree in depth-first order (-1: end)
  int pad;
};
// BVH2D::build(fwd, leaf_size, max_depth) of the (n+1)^2-vertex quad grid with
// SoA positions d_xy (device) on stream s: *nodes_out (num_nodes records) and
// *faces_out (n^2 face slots) are new device allocations owned by the caller.
void bvh_build_device(int n, const double* d_xy, int leaf_size, int max_depth, hipStream_t s, BNode** nodes_out,
                      int** faces_out, int* num_nodes);

}  // namespace otm
