// Internal declarations shared by the HIP translation units of libotmap_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

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

// Per-grid device buffers.
struct Fields {
  int n = 0;
  long F = 0, V = 0;
  double* rho = nullptr;     // adjusted density (F)
  double* psi = nullptr;     // potential (F)
  double* g0 = nullptr;      // cached vertex gradients of psi (2V, SoA)
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
  double* site(int s) const { return partials + long(s) * SITE_VALUES * pslots; }
};

// Stencil kernel launchers (stencil.hip)
void launch_fill(double* p, long count, double v, hipStream_t s);
void launch_adjust_density(const double* in, const Fields& f, double max_ratio, hipStream_t s);
void launch_initial_residual(const Fields& f, int max_iter, double threshold, int cj, hipStream_t s);
void launch_beta(const Fields& f, hipStream_t s);       // k_beta + scalar kernel
void launch_dir_coef(const Fields& f, hipStream_t s);   // k_dir_coef + scalar kernel (line search)
void launch_update(const Fields& f, hipStream_t s);
// number of workgroups of the final DCT stage (its |d_hat|^2 partials go to SITE_DNORM)
int dnorm_blocks(int n, bool mixed);
void launch_forward_positions(const double* psi, double* xy, int n, hipStream_t s);
// component kernels (single evaluations)
void launch_vertex_gradients(const double* psi, double* G, int n, hipStream_t s);
void launch_residual_only(const double* psi, const double* rho, double* r, double* partials, int n, hipStream_t s);
void launch_params_1d(const double* g0, const double* dir, double* a, double* b, int n, hipStream_t s);

// DCT Laplacian solve (dct.hip): d_hat = M^+ (r - S e0), S = state->sum_r
struct DctPlan {
  int n = 0;
  double* C = nullptr;    // orthonormal DCT-II matrix, row-major C[k][i]
  double* Ct = nullptr;   // its transpose
  double* sq = nullptr;   // sin^2(pi m / (2n)), m = 0..2n
  // even/odd folded form (n % 128 == 0): with h = n/2,
  //   P[r][i] = C[2r][i] (r < h), C[2(r-h)+1][i] (r >= h), i < h   (n x h)
  //   Pt = P^T                                                      (h x n)
  bool fold = false;
  double* P = nullptr;
  double* Pt = nullptr;
  int cfg[4] = {0, 0, 0, 0};  // tile shape per folded stage, picked by dct_plan_tune
};
void dct_plan_create(DctPlan& p, int n);
void dct_plan_destroy(DctPlan& p);
void launch_dct_solve(const DctPlan& p, const Fields& f, const double* r, double* out, hipStream_t s);
void launch_dct_gemm_stage(const DctPlan& p, const Fields& f, hipStream_t s);
// times every candidate tile shape of each folded stage on the handle's stream
// (once per device and n, cached for the process) and stores the fastest
void dct_plan_tune(DctPlan& p, int device, const Fields& f, hipStream_t s);
int dct_fold_configs();

// Mixed-precision Laplacian solve (dct_mixed.hip): 3-term BF16 split GEMMs on
// MFMA + FP64 refinement.  Requires n % 64 == 0.
struct MixedPlan {
  int n = 0;
  float* C32 = nullptr;   // C   [k][i] in FP32 (split into BF16 terms on the fly)
  float* Ct32 = nullptr;  // C^T [i][k]
  float* work = nullptr;  // 2 intermediates (FP32)
  float* rhs = nullptr;   // right-hand side of the current transform (FP32)
  double* x0 = nullptr;   // unrefined solution (FP64)
};
bool mixed_supported(int n);
void mixed_plan_create(MixedPlan& mp, const DctPlan& p);
void mixed_plan_destroy(MixedPlan& mp);
void launch_mixed_solve(const MixedPlan& mp, const DctPlan& p, const Fields& f, const double* r, double* out,
                        int refine, hipStream_t s);
void launch_mixed_gemm_stage(const MixedPlan& mp, const DctPlan& p, hipStream_t s);

}  // namespace otm
