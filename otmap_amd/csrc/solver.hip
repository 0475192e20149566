// Solver handle: device buffers for one grid size, the DCT plan, and the
// host-side orchestration of GridBasedTransportSolver::init()/solve()
// (otlib/otsolver_2dgrid.cpp:75-245).
//
// One Newton iteration is eight kernels on the handle's stream with the
// default fast-DCT transform (nine with the GEMM transform), each launched
// over the handle's one-job table:
//   3 fast-DCT passes, or 4 DCT GEMM stages (d_hat = M^+ (r - S e0))  otsolver_2dgrid.cpp:174-183
//   k_beta + k_scalar_beta (conjugate-Jacobian beta)                  :191-200, :248-261
//   k_dir_coef + k_scalar_step (d, quartic coefficients, alpha)       :467-598
//   k_update (psi, r, Jd)                                              :599-605
// All scalar decisions are taken on the device; the host only checks the
// stop flag once per chunk of iterations, and kernels enqueued past the stop
// return immediately.

#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/otmap_capi.h"
#include "otm_internal.h"

namespace otm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace otm

using namespace otm;

struct otm_solver {
  int device = 0;
  int verbose = 1;
  int n = 0;
  Fields f;                    // the grid's buffers (also js.host[0].f)
  JobSet js;                   // the one-job table every kernel of a solve runs over
  Driver drv;                  // stream, captured iterations, events
  DctPlan plan;
  MixedPlan mplan;             // created on the first mixed-precision solve
  double* d_in = nullptr;      // staging of host-API inputs (n*n)
  double* d_xy = nullptr;      // staging of host-API vertex outputs (2V)
  SolveState* h_state = nullptr;  // pinned
  int last_iters = 0;
};

namespace {

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <class T>
void dalloc(T*& p, size_t count) {
  OTM_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
}

void release_fields(otm_solver* s) {
  driver_invalidate(s->drv);
  fields_destroy(s->f);
  dfree(s->d_in); dfree(s->d_xy);
  jobset_destroy(s->js);
  mixed_plan_destroy(s->mplan);
  dct_plan_destroy(s->plan);
  s->n = 0;
}

void allocate(otm_solver* s, int n) {
  fields_create(s->f, n, 1024);
  dalloc(s->d_in, s->f.F);
  dalloc(s->d_xy, 2 * s->f.V);
  dct_plan_create(s->plan, n);
  Job J;
  J.f = s->f;
  J.P = s->plan.P;
  J.Pt = s->plan.Pt;
  J.sq = s->plan.sq;
  J.fft = s->plan.fft;
  jobset_create(s->js, {J});
  dct_tune(s->js, s->device, s->drv.stream);
  s->n = n;
}

// mixed-precision buffers and BF16 factor planes, on first use (the tile
// configurations are tuned with the job on the mixed transform)
void ensure_mixed(otm_solver* s) {
  if (s->f.x0) return;
  fields_enable_mixed(s->f);
  mixed_plan_create(s->mplan, s->plan);
  Job& J = s->js.host[0];
  J.f = s->f;
  J.Pp = s->mplan.Pp;
  J.Ptp = s->mplan.Ptp;
  J.ck = s->mplan.ck;
  jobset_upload(s->js, s->drv.stream);
  const std::vector<int> route = s->js.jroute;
  jobset_route(s->js, {1});
  mixed_tune(s->js, s->device, s->drv.stream);
  jobset_route(s->js, route);
  driver_invalidate(s->drv);
}

int fail_hip(const HipError& e) {
  set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
  return OTM_ERR_HIP;
}

#define OTM_GUARD_BEGIN try {
#define OTM_GUARD_END                         \
  }                                           \
  catch (const HipError& e) {                 \
    return fail_hip(e);                       \
  }                                           \
  catch (const std::bad_alloc&) {             \
    set_error("host allocation failed");      \
    return OTM_ERR_ALLOC;                     \
  }

int check_ready(otm_solver* s) {
  if (!s) { set_error("null solver handle"); return OTM_ERR_ARG; }
  if (s->n <= 0) { set_error("solver not initialised: call otm_solver_init(n) first"); return OTM_ERR_STATE; }
  return OTM_OK;
}

int solve_device(otm_solver* s, const double* d_density, const otm_options* opt_in, double* d_fwd, double* d_adj,
                 otm_stats* stats) {
  otm_options opt;
  otm_options_init(&opt);
  if (opt_in) opt = *opt_in;
  if (opt.max_iter < 0) opt.max_iter = 0;
  if (opt.precision != OTM_PRECISION_FP64 && opt.precision != OTM_PRECISION_MIXED) {
    set_error("unknown precision mode");
    return OTM_ERR_ARG;
  }
  if (opt.transform != OTM_TRANSFORM_GEMM && opt.transform != OTM_TRANSFORM_FFT) {
    set_error("unknown transform");
    return OTM_ERR_ARG;
  }
  int route = 0;
  if (opt.precision == OTM_PRECISION_MIXED) {
    if (opt.refine_steps < 0 || opt.refine_steps > 8) {
      set_error("refine_steps must be in [0, 8]");
      return OTM_ERR_ARG;
    }
    route = mixed_route(s->n, opt.mixed_min_n);
    if (route) ensure_mixed(s);
  }
  if (route == 0) route = fp64_route(s->n, opt.transform);
  jobset_route(s->js, {route});
  if (fields_grow_trace(s->f, trace_cap_for(opt.max_iter))) driver_invalidate(s->drv);
  Job& J = s->js.host[0];
  J.f = s->f;
  J.dens_in = d_density;
  J.fwd_out = d_fwd;
  J.adj_out = d_adj;
  float ms = 0.f;
  if (int rc = driver_solve(s->drv, s->js, &s->plan, opt, &ms)) return rc;
  OTM_HIP(hipMemcpy(s->h_state, s->f.state, sizeof(SolveState), hipMemcpyDeviceToHost));
  const SolveState& hs = *s->h_state;
  s->last_iters = hs.it;
  if (stats) {
    stats->iterations = hs.it;
    stats->exit_reason = hs.exit_reason;
    stats->initial_residual = hs.e0;
    stats->final_residual = hs.e;
    stats->ms_solve = ms;
  }
  return OTM_OK;
}

}  // namespace

extern "C" {

void otm_options_init(otm_options* opt) {
  if (!opt) return;
  opt->beta = OTM_BETA_CONJUGATE_JACOBIAN;
  opt->max_iter = 1000;
  opt->threshold = 1e-7;
  opt->max_ratio = DBL_MAX;
  opt->precision = OTM_PRECISION_FP64;
  opt->refine_steps = 1;
  opt->mixed_min_n = 1024;
  opt->transform = OTM_TRANSFORM_FFT;
}

const char* otm_last_error(void) { return g_last_error.c_str(); }

int otm_solver_create(otm_solver** out, int device) {
  if (!out) { set_error("null output pointer"); return OTM_ERR_ARG; }
  *out = nullptr;
  OTM_GUARD_BEGIN
  int count = 0;
  OTM_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) { set_error("invalid HIP device index"); return OTM_ERR_ARG; }
  OTM_HIP(hipSetDevice(device));
  auto* s = new otm_solver;
  s->device = device;
  driver_create(s->drv);
  OTM_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_state), sizeof(SolveState), hipHostMallocDefault));
  *out = s;
  return OTM_OK;
  OTM_GUARD_END
}

int otm_solver_destroy(otm_solver* s) {
  if (!s) return OTM_OK;
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  if (s->drv.stream) OTM_HIP(hipStreamSynchronize(s->drv.stream));
  release_fields(s);
  if (s->h_state) (void)hipHostFree(s->h_state);
  driver_destroy(s->drv);
  delete s;
  return OTM_OK;
  OTM_GUARD_END
}

int otm_solver_set_verbose_level(otm_solver* s, int level) {
  if (!s) { set_error("null solver handle"); return OTM_ERR_ARG; }
  s->verbose = level;
  return OTM_OK;
}

int otm_solver_init(otm_solver* s, int n) {
  if (!s) { set_error("null solver handle"); return OTM_ERR_ARG; }
  if (n < 1 || n > 16384) { set_error("grid size must be in [1, 16384]"); return OTM_ERR_ARG; }
  if (s->n == n) return OTM_OK;  // already set up (otsolver_2dgrid.cpp:79-83)
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  release_fields(s);
  allocate(s, n);
  return OTM_OK;
  OTM_GUARD_END
}

int otm_solver_grid_size(const otm_solver* s) { return s ? s->n : 0; }

int otm_solver_gemm_config(const otm_solver* s, int* cfg) {
  if (int rc = check_ready(const_cast<otm_solver*>(s))) return rc;
  if (!cfg) { set_error("null config array"); return OTM_ERR_ARG; }
  for (int st = 0; st < 4; ++st) cfg[st] = s->js.cfg[st];
  return s->js.fold ? NFOLD : 0;
}

int otm_solver_set_gemm_config(otm_solver* s, int stage, int config) {
  if (int rc = check_ready(s)) return rc;
  if (!s->js.fold) { set_error("the DCT of this grid size does not fold (n % 64 != 0)"); return OTM_ERR_STATE; }
  if (stage < 0 || stage > 3 || config < 0 || config >= NFOLD) {
    set_error("stage must be 0..3 and config 0..otm_solver_gemm_config()-1");
    return OTM_ERR_ARG;
  }
  if (!fold_cfg_valid(s->n, config)) { set_error("tile shape does not divide n/2"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  s->js.jcfg[0][stage] = config;
  jobset_assign(s->js);
  return OTM_OK;
  OTM_GUARD_END
}

int otm_solver_mixed_config(const otm_solver* s, int* cfg) {
  if (int rc = check_ready(const_cast<otm_solver*>(s))) return rc;
  if (!cfg) { set_error("null config array"); return OTM_ERR_ARG; }
  for (int st = 0; st < 4; ++st) cfg[st] = s->js.mcfg[st];
  return s->f.x0 ? NMIXED : 0;
}

int otm_solver_set_mixed_config(otm_solver* s, int stage, int config) {
  if (int rc = check_ready(s)) return rc;
  if (!s->f.x0) { set_error("no mixed solve has run on this solver"); return OTM_ERR_STATE; }
  if (stage < 0 || stage > 3 || config < 0 || config >= NMIXED) {
    set_error("stage must be 0..3 and config 0..otm_solver_mixed_config()-1");
    return OTM_ERR_ARG;
  }
  if (!mixed_cfg_valid(stage, s->n, config)) { set_error("tile width does not divide the stage"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  s->js.jmcfg[0][stage] = config;
  jobset_assign(s->js);
  return OTM_OK;
  OTM_GUARD_END
}

int otm_solve_device(otm_solver* s, const double* d_density, const otm_options* opt, double* d_fwd_xy,
                     double* d_adjusted_density, otm_stats* stats) {
  if (int rc = check_ready(s)) return rc;
  if (!d_density || !d_fwd_xy) { set_error("null density or output pointer"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  return solve_device(s, d_density, opt, d_fwd_xy, d_adjusted_density, stats);
  OTM_GUARD_END
}

int otm_solve(otm_solver* s, const double* density, const otm_options* opt, double* fwd_xy, double* adjusted_density,
              otm_stats* stats) {
  if (int rc = check_ready(s)) return rc;
  if (!density || !fwd_xy) { set_error("null density or output pointer"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  const Fields& f = s->f;
  OTM_HIP(hipMemcpyAsync(s->d_in, density, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  int rc = solve_device(s, s->d_in, opt, s->d_xy, nullptr, stats);
  if (rc) return rc;
  OTM_HIP(hipMemcpyAsync(fwd_xy, s->d_xy, 2 * f.V * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  if (adjusted_density)
    OTM_HIP(hipMemcpyAsync(adjusted_density, f.rho, f.F * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  return OTM_OK;
  OTM_GUARD_END
}

int otm_solver_trace(const otm_solver* s, double* residual, double* alpha, double* beta, int cap) {
  if (!s || s->n <= 0) { set_error("solver not initialised"); return OTM_ERR_STATE; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  const int k = std::min(std::min(cap, s->last_iters), s->f.trace_cap);
  if (k <= 0) return 0;
  const int tc = s->f.trace_cap;
  if (residual) OTM_HIP(hipMemcpy(residual, s->f.trace, k * sizeof(double), hipMemcpyDeviceToHost));
  if (alpha) OTM_HIP(hipMemcpy(alpha, s->f.trace + tc, k * sizeof(double), hipMemcpyDeviceToHost));
  if (beta) OTM_HIP(hipMemcpy(beta, s->f.trace + 2 * tc, k * sizeof(double), hipMemcpyDeviceToHost));
  return k;
  OTM_GUARD_END
}

int otm_solver_profile_kernel(otm_solver* s, int kernel, int reps, double* ms_per_launch) {
  if (int rc = check_ready(s)) return rc;
  if (reps < 1 || !ms_per_launch || kernel < OTM_KERNEL_DCT_GEMM || kernel > OTM_KERNEL_DIR_COEF_PASS) {
    set_error("invalid profile arguments");
    return OTM_ERR_ARG;
  }
  const bool mixed = kernel == OTM_KERNEL_MIXED_GEMM || kernel == OTM_KERNEL_MIXED_SOLVE;
  if (mixed && !mixed_supported(s->n)) {
    set_error("mixed precision needs a grid size that is a multiple of 64");
    return OTM_ERR_ARG;
  }
  const bool fft = kernel >= OTM_KERNEL_FFT_SOLVE && kernel <= OTM_KERNEL_FFT_ROWS_INV;
  if (fft && !fft_supported(s->n)) {
    set_error("the fast-DCT transform does not support this grid size");
    return OTM_ERR_ARG;
  }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  if (mixed) ensure_mixed(s);
  // A scratch state keeps every gate open: iteration 1, conjugate-Jacobian
  // beta, update pending.  Intermediate buffers of the last solve are reused
  // as inputs and overwritten.
  SolveState* scratch = nullptr;
  OTM_HIP(hipMalloc(&scratch, sizeof(SolveState)));
  Job J = s->js.host[0];
  J.f.state = scratch;
  Fields& f = J.f;
  JobSet pj;
  jobset_create(pj, {J});
  pj.jcfg = s->js.jcfg;
  pj.jmcfg = s->js.jmcfg;
  pj.jroute = {mixed ? 1 : fft ? 2 : 0};
  pj.routed_mixed = mixed ? 1 : 0;
  pj.routed_fft = fft ? 1 : 0;
  jobset_assign(pj);
  SolveState tpl;
  std::memset(&tpl, 0, sizeof(tpl));
  tpl.it = 1; tpl.cj = 1; tpl.max_iter = 1 << 30; tpl.threshold = 0.; tpl.alpha = 0.5; tpl.beta = 0.1;
  tpl.e = 1.; tpl.update = 1; tpl.use_prev = 1;
  hipEvent_t a, b;
  OTM_HIP(hipEventCreate(&a));
  OTM_HIP(hipEventCreate(&b));
  // The gates stay open across repeated launches (the scalar kernels only
  // advance the iteration count), so every kernel runs back-to-back between
  // one event pair: the figure is the kernel duration with launch latency
  // overlapped, as rocprof reports it.  BETA and DIR_COEF include their
  // single-workgroup scalar kernel; DIR_COEF_PASS is the line-search pass alone.
  if (kernel == OTM_KERNEL_DCT_SOLVE || kernel == OTM_KERNEL_MIXED_SOLVE || fft)
    tpl.update = 0;
  OTM_HIP(hipMemcpyAsync(scratch, &tpl, sizeof(tpl), hipMemcpyHostToDevice, s->drv.stream));
  auto launch = [&]() {
    switch (kernel) {
      case OTM_KERNEL_DCT_GEMM: launch_dct_gemm_stage(pj, &s->plan, s->drv.stream); break;
      case OTM_KERNEL_BETA: launch_beta(pj, s->drv.stream); break;
      case OTM_KERNEL_DIR_COEF: launch_dir_coef(pj, s->drv.stream); break;
      case OTM_KERNEL_UPDATE: launch_update(pj, s->drv.stream); break;
      case OTM_KERNEL_DCT_SOLVE: launch_dct_solve(pj, &s->plan, s->drv.stream); break;
      case OTM_KERNEL_MIXED_GEMM: launch_mixed_gemm_stage(pj, s->drv.stream); break;
      case OTM_KERNEL_FFT_SOLVE: launch_fft_solve(pj, true, s->drv.stream); break;
      case OTM_KERNEL_FFT_ROWS_FWD: launch_fft_pass(pj, 0, true, s->drv.stream); break;
      case OTM_KERNEL_FFT_COLS: launch_fft_pass(pj, 1, true, s->drv.stream); break;
      case OTM_KERNEL_FFT_ROWS_INV: launch_fft_pass(pj, 2, true, s->drv.stream); break;
      case OTM_KERNEL_DIR_COEF_PASS: launch_dir_coef_pass(pj, s->drv.stream); break;
      default: launch_mixed_solve(pj, 1, true, s->drv.stream); break;
    }
  };
  launch();  // warm-up
  OTM_HIP(hipEventRecord(a, s->drv.stream));
  for (int k = 0; k < reps; ++k) launch();
  OTM_HIP(hipEventRecord(b, s->drv.stream));
  OTM_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  OTM_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  jobset_destroy(pj);
  (void)hipFree(scratch);
  *ms_per_launch = double(ms) / reps;
  return OTM_OK;
  OTM_GUARD_END
}

int otm_vertex_gradients(otm_solver* s, const double* psi, double* grads) {
  if (int rc = check_ready(s)) return rc;
  if (!psi || !grads) { set_error("null pointer"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  const Fields& f = s->f;
  OTM_HIP(hipMemcpyAsync(s->d_in, psi, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  launch_vertex_gradients(s->d_in, s->d_xy, s->n, s->drv.stream);
  OTM_HIP(hipMemcpyAsync(grads, s->d_xy, 2 * f.V * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  return OTM_OK;
  OTM_GUARD_END
}

int otm_residual(otm_solver* s, const double* adjusted_density, const double* psi, double* r, double* e) {
  if (int rc = check_ready(s)) return rc;
  if (!adjusted_density || !psi || !r) { set_error("null pointer"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  const Fields& f = s->f;
  const int nb = num_tiles(s->n);
  OTM_HIP(hipMemcpyAsync(f.rho, adjusted_density, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  OTM_HIP(hipMemcpyAsync(f.psi, psi, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  launch_residual_only(f.psi, f.rho, f.r, f.partials, s->n, s->drv.stream);
  OTM_HIP(hipMemcpyAsync(r, f.r, f.F * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  std::vector<double> part(nb);
  OTM_HIP(hipMemcpyAsync(part.data(), f.partials, nb * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  double acc = 0.;
  for (double p : part) acc += p;
  if (e) *e = acc / (1.0 / (double(s->n) * double(s->n)));
  return OTM_OK;
  OTM_GUARD_END
}

int otm_laplacian_solve(otm_solver* s, const double* r, double* d_hat) {
  if (int rc = check_ready(s)) return rc;
  if (!r || !d_hat) { set_error("null pointer"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  jobset_route(s->js, {0});
  const Fields& f = s->f;
  double sum = 0.;
  for (long k = 0; k < f.F; ++k) sum += r[k];
  SolveState st;
  std::memset(&st, 0, sizeof(st));
  st.sum_r = sum;
  OTM_HIP(hipMemcpyAsync(f.state, &st, sizeof(st), hipMemcpyHostToDevice, s->drv.stream));
  OTM_HIP(hipMemcpyAsync(f.r, r, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  launch_dct_solve(s->js, &s->plan, s->drv.stream);
  OTM_HIP(hipMemcpyAsync(d_hat, f.dhat, f.F * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  return OTM_OK;
  OTM_GUARD_END
}

int otm_laplacian_solve_mixed(otm_solver* s, const double* r, double* d_hat, int refine_steps) {
  if (int rc = check_ready(s)) return rc;
  if (!r || !d_hat) { set_error("null pointer"); return OTM_ERR_ARG; }
  if (!mixed_supported(s->n)) { set_error("mixed precision needs a grid size that is a multiple of 64"); return OTM_ERR_ARG; }
  if (refine_steps < 0 || refine_steps > 8) { set_error("refine_steps must be in [0, 8]"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  ensure_mixed(s);
  jobset_route(s->js, {1});
  const Fields& f = s->f;
  double sum = 0.;
  for (long k = 0; k < f.F; ++k) sum += r[k];
  SolveState st;
  std::memset(&st, 0, sizeof(st));
  st.sum_r = sum;
  OTM_HIP(hipMemcpyAsync(f.state, &st, sizeof(st), hipMemcpyHostToDevice, s->drv.stream));
  OTM_HIP(hipMemcpyAsync(f.r, r, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  launch_mixed_solve(s->js, refine_steps, true, s->drv.stream);
  OTM_HIP(hipMemcpyAsync(d_hat, f.dhat, f.F * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  return OTM_OK;
  OTM_GUARD_END
}

int otm_laplacian_solve_fft(otm_solver* s, const double* r, double* d_hat) {
  if (int rc = check_ready(s)) return rc;
  if (!r || !d_hat) { set_error("null pointer"); return OTM_ERR_ARG; }
  if (!fft_supported(s->n)) {
    set_error("the fast-DCT transform needs n = 2^a 3^b with 8 <= n <= 2048 and n % 4 == 0");
    return OTM_ERR_ARG;
  }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  jobset_route(s->js, {2});
  const Fields& f = s->f;
  double sum = 0.;
  for (long k = 0; k < f.F; ++k) sum += r[k];
  SolveState st;
  std::memset(&st, 0, sizeof(st));
  st.sum_r = sum;
  OTM_HIP(hipMemcpyAsync(f.state, &st, sizeof(st), hipMemcpyHostToDevice, s->drv.stream));
  OTM_HIP(hipMemcpyAsync(f.r, r, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  launch_fft_solve(s->js, true, s->drv.stream);
  OTM_HIP(hipMemcpyAsync(d_hat, f.dhat, f.F * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  return OTM_OK;
  OTM_GUARD_END
}

int otm_line_search_coefficients(otm_solver* s, const double* g0, const double* dir, double* a, double* b) {
  if (int rc = check_ready(s)) return rc;
  if (!g0 || !dir || !a || !b) { set_error("null pointer"); return OTM_ERR_ARG; }
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(s->device));
  const Fields& f = s->f;
  OTM_HIP(hipMemcpyAsync(f.g0, g0, 2 * f.V * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  OTM_HIP(hipMemcpyAsync(f.dhat, dir, f.F * sizeof(double), hipMemcpyHostToDevice, s->drv.stream));
  launch_params_1d(f.g0, f.dhat, f.a, f.b, s->n, s->drv.stream);
  OTM_HIP(hipMemcpyAsync(a, f.a, f.F * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipMemcpyAsync(b, f.b, f.F * sizeof(double), hipMemcpyDeviceToHost, s->drv.stream));
  OTM_HIP(hipStreamSynchronize(s->drv.stream));
  return OTM_OK;
  OTM_GUARD_END
}

int otm_line_search_steps(int device, const double* z, int count, double* alpha, double* emin) {
  if (count < 0 || (count > 0 && (!z || !alpha || !emin))) { set_error("null pointer or negative count"); return OTM_ERR_ARG; }
  if (count == 0) return OTM_OK;
  OTM_GUARD_BEGIN
  OTM_HIP(hipSetDevice(device));
  double *dz = nullptr, *da = nullptr, *de = nullptr;
  OTM_HIP(hipMalloc(&dz, 5 * size_t(count) * sizeof(double)));
  hipError_t rc = hipMalloc(&da, size_t(count) * sizeof(double));
  if (rc == hipSuccess) rc = hipMalloc(&de, size_t(count) * sizeof(double));
  if (rc == hipSuccess) rc = hipMemcpy(dz, z, 5 * size_t(count) * sizeof(double), hipMemcpyHostToDevice);
  if (rc == hipSuccess) {
    launch_pick_steps(dz, count, da, de, nullptr);
    rc = hipMemcpy(alpha, da, size_t(count) * sizeof(double), hipMemcpyDeviceToHost);
  }
  if (rc == hipSuccess) rc = hipMemcpy(emin, de, size_t(count) * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dz);
  (void)hipFree(da);
  (void)hipFree(de);
  OTM_HIP(rc);
  return OTM_OK;
  OTM_GUARD_END
}

}  // extern "C"
