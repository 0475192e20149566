// Batched solves: many independent densities of uneven grid sizes solved
// together.  The reference solves one density per GridBasedTransportSolver
// (otlib/otsolver_2dgrid.cpp:106-245); here every grid of the batch is one job
// of a job table, so each step of a Newton iteration is one launch for the
// whole batch, the iterations are captured in a HIP graph, and the DCT tile
// shapes are picked by the autotuner per grid size (each size class runs its
// own tile configuration, one launch per configuration in use).  Each job
// runs exactly the single-grid computation (same kernels, same operation
// order), so a batched result is bit-identical to the same density solved
// alone; jobs that converge early skip the remaining iterations.  Both
// precision modes run batched (mixed: the folded BF16x3 transform with FP64
// refinement for the grids from mixed_min_n up, the FP64 transform for the
// others, in the same launches).

#include <cstring>
#include <map>
#include <vector>

#include "otm_internal.h"

using namespace otm;

struct otm_batch {
  int device = 0;
  int count = 0;
  std::vector<int> n;
  std::vector<Fields> f;
  std::map<int, DctPlan> plans;     // one plan per distinct grid size
  std::map<int, MixedPlan> mplans;  // BF16 factor planes, on the first mixed solve
  JobSet js;
  Driver drv;
  std::vector<double*> d_in, d_xy;  // staging of host-API buffers
  int trace_cap = 0;
  std::vector<int> mixed_tuned_route;  // routes for which the mixed tile widths were picked
};

namespace {

int fail_hip(const HipError& e) {
  set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
  return OTM_ERR_HIP;
}

void batch_release(otm_batch* b) {
  driver_destroy(b->drv);
  jobset_destroy(b->js);
  for (Fields& f : b->f) fields_destroy(f);
  for (auto& kv : b->plans) dct_plan_destroy(kv.second);
  for (auto& kv : b->mplans) mixed_plan_destroy(kv.second);
  b->mplans.clear();
  for (double* p : b->d_in) if (p) (void)hipFree(p);
  for (double* p : b->d_xy) if (p) (void)hipFree(p);
  b->f.clear();
  b->plans.clear();
  b->d_in.clear();
  b->d_xy.clear();
}

void build_table(otm_batch* b) {
  std::vector<Job> jobs(b->count);
  for (int p = 0; p < b->count; ++p) {
    const DctPlan& pl = b->plans.at(b->n[p]);
    jobs[p].f = b->f[p];
    jobs[p].P = pl.P;
    jobs[p].Pt = pl.Pt;
    jobs[p].sq = pl.sq;
    jobs[p].fft = pl.fft;
    auto it = b->mplans.find(b->n[p]);
    if (it != b->mplans.end()) {
      jobs[p].Pp = it->second.Pp;
      jobs[p].Ptp = it->second.Ptp;
      jobs[p].ck = it->second.ck;
    }
  }
  // the tile assignment and the routes survive a rebuild (same jobs, same order)
  const bool had = b->js.jobs != nullptr;
  const auto jcfg = b->js.jcfg;
  const auto jmcfg = b->js.jmcfg;
  const auto jroute = b->js.jroute;
  const int routed = b->js.routed_mixed, routed_fft = b->js.routed_fft;
  jobset_destroy(b->js);
  jobset_create(b->js, jobs);
  if (had) {
    b->js.jcfg = jcfg;
    b->js.jmcfg = jmcfg;
    b->js.jroute = jroute;
    b->js.routed_mixed = routed;
    b->js.routed_fft = routed_fft;
    jobset_assign(b->js);
  }
}

int solve_batch(otm_batch* b, const double* const* d_density, const otm_options* opt_in, double* const* d_fwd,
                double* const* d_adj, otm_stats* stats) {
  otm_options opt;
  otm_options_init(&opt);
  if (opt_in) opt = *opt_in;
  if (opt.max_iter < 0) opt.max_iter = 0;
  if (opt.precision != OTM_PRECISION_FP64 && opt.precision != OTM_PRECISION_MIXED) {
    set_error("unknown precision mode");
    return OTM_ERR_ARG;
  }
  if (opt.precision == OTM_PRECISION_MIXED && (opt.refine_steps < 0 || opt.refine_steps > 8)) {
    set_error("refine_steps must be in [0, 8]");
    return OTM_ERR_ARG;
  }
  bool rebuild = false;
  const int cap = trace_cap_for(opt.max_iter);
  if (cap > b->trace_cap) {
    for (Fields& f : b->f) fields_grow_trace(f, cap);
    b->trace_cap = cap;
    rebuild = true;
  }
  if (opt.transform != OTM_TRANSFORM_GEMM && opt.transform != OTM_TRANSFORM_FFT) {
    set_error("unknown transform");
    return OTM_ERR_ARG;
  }
  // each grid's transform: mixed precision runs the BF16x3 transform on the
  // grids from mixed_min_n up; the others run the FP64 transform, as GEMMs or
  // as fast DCTs (opt.transform, where the size allows)
  std::vector<int> route(b->count, 0);
  for (int p = 0; p < b->count; ++p) {
    if (opt.precision == OTM_PRECISION_MIXED) route[p] = mixed_route(b->n[p], opt.mixed_min_n);
    if (route[p] == 0) route[p] = fp64_route(b->n[p], opt.transform);
  }
  for (int p = 0; p < b->count; ++p)
    if (route[p] == 1 && !b->f[p].x0) {  // mixed buffers and BF16 factor planes on first use
      fields_enable_mixed(b->f[p]);
      if (!b->mplans.count(b->n[p])) mixed_plan_create(b->mplans[b->n[p]], b->plans.at(b->n[p]));
      rebuild = true;
    }
  if (rebuild) {
    build_table(b);
    driver_invalidate(b->drv);
  }
  jobset_route(b->js, route);
  if (b->js.routed_mixed > 0 && b->mixed_tuned_route != route) {
    mixed_tune(b->js, b->device, b->drv.stream);
    b->mixed_tuned_route = route;
  }
  for (int p = 0; p < b->count; ++p) {
    Job& J = b->js.host[p];
    J.dens_in = d_density[p];
    J.fwd_out = d_fwd[p];
    J.adj_out = d_adj ? d_adj[p] : nullptr;
  }
  float ms = 0.f;
  if (int rc = driver_solve(b->drv, b->js, nullptr, opt, &ms)) return rc;
  if (stats) {
    for (int p = 0; p < b->count; ++p) {
      SolveState hs;
      OTM_HIP(hipMemcpy(&hs, b->f[p].state, sizeof(SolveState), hipMemcpyDeviceToHost));
      stats[p].iterations = hs.it;
      stats[p].exit_reason = hs.exit_reason;
      stats[p].initial_residual = hs.e0;
      stats[p].final_residual = hs.e;
      stats[p].ms_solve = ms;
    }
  }
  return OTM_OK;
}

}  // namespace

extern "C" {

int otm_batch_create(otm_batch** out, int device, int count, const int* n) {
  if (!out || !n || count < 1) { set_error("invalid batch arguments"); return OTM_ERR_ARG; }
  *out = nullptr;
  for (int p = 0; p < count; ++p)
    if (n[p] < 64 || n[p] > 16384 || n[p] % 64 != 0) {
      set_error("batched grid sizes must be multiples of 64 in [64, 16384]");
      return OTM_ERR_ARG;
    }
  auto* b = new otm_batch;
  try {
    int ndev = 0;
    OTM_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
      delete b;
      set_error("invalid HIP device index");
      return OTM_ERR_ARG;
    }
    OTM_HIP(hipSetDevice(device));
    b->device = device;
    b->count = count;
    b->n.assign(n, n + count);
    driver_create(b->drv);
    b->trace_cap = 1024;
    b->f.resize(count);
    for (int p = 0; p < count; ++p) {
      fields_create(b->f[p], n[p], b->trace_cap);
      if (!b->plans.count(n[p])) dct_plan_create(b->plans[n[p]], n[p]);
    }
    build_table(b);
    dct_tune(b->js, device, b->drv.stream);
    *out = b;
    return OTM_OK;
  } catch (const HipError& e) {
    batch_release(b);
    delete b;
    return fail_hip(e);
  }
}

int otm_batch_destroy(otm_batch* b) {
  if (!b) return OTM_OK;
  try {
    OTM_HIP(hipSetDevice(b->device));
    batch_release(b);
    delete b;
    return OTM_OK;
  } catch (const HipError& e) {
    return fail_hip(e);
  }
}

int otm_batch_size(const otm_batch* b) { return b ? b->count : 0; }

int otm_batch_gemm_config(const otm_batch* b, int job, int* cfg) {
  if (!b || !cfg) { set_error("null argument"); return OTM_ERR_ARG; }
  if (job < -1 || job >= b->count) { set_error("job index out of range"); return OTM_ERR_ARG; }
  for (int st = 0; st < 4; ++st) cfg[st] = job < 0 ? b->js.cfg[st] : b->js.jcfg[job][st];
  return NFOLD;
}

int otm_batch_mixed_config(const otm_batch* b, int job, int* cfg) {
  if (!b || !cfg) { set_error("null argument"); return OTM_ERR_ARG; }
  if (job < -1 || job >= b->count) { set_error("job index out of range"); return OTM_ERR_ARG; }
  for (int st = 0; st < 4; ++st) cfg[st] = job < 0 ? b->js.mcfg[st] : b->js.jmcfg[job][st];
  return b->mixed_tuned_route.empty() ? 0 : NMIXED;
}

int otm_batch_solve_device(otm_batch* b, const double* const* d_density, const otm_options* opt,
                           double* const* d_fwd_xy, double* const* d_adjusted_density, otm_stats* stats) {
  if (!b || !d_density || !d_fwd_xy) { set_error("null argument"); return OTM_ERR_ARG; }
  for (int p = 0; p < b->count; ++p)
    if (!d_density[p] || !d_fwd_xy[p]) { set_error("null density or output pointer"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(b->device));
    return solve_batch(b, d_density, opt, d_fwd_xy, d_adjusted_density, stats);
  } catch (const HipError& e) {
    return fail_hip(e);
  }
}

int otm_batch_solve(otm_batch* b, const double* const* density, const otm_options* opt, double* const* fwd_xy,
                    double* const* adjusted_density, otm_stats* stats) {
  if (!b || !density || !fwd_xy) { set_error("null argument"); return OTM_ERR_ARG; }
  for (int p = 0; p < b->count; ++p)
    if (!density[p] || !fwd_xy[p]) { set_error("null density or output pointer"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(b->device));
    hipStream_t s = b->drv.stream;
    if (b->d_in.empty()) {
      b->d_in.assign(b->count, nullptr);
      b->d_xy.assign(b->count, nullptr);
      for (int p = 0; p < b->count; ++p) {
        OTM_HIP(hipMalloc(&b->d_in[p], b->f[p].F * sizeof(double)));
        OTM_HIP(hipMalloc(&b->d_xy[p], 2 * b->f[p].V * sizeof(double)));
      }
    }
    for (int p = 0; p < b->count; ++p)
      OTM_HIP(hipMemcpyAsync(b->d_in[p], density[p], b->f[p].F * sizeof(double), hipMemcpyHostToDevice, s));
    std::vector<const double*> din(b->d_in.begin(), b->d_in.end());
    std::vector<double*> dadj(b->count, nullptr);
    for (int p = 0; p < b->count; ++p) dadj[p] = (adjusted_density && adjusted_density[p]) ? b->f[p].rho : nullptr;
    // adjusted densities are read back from the grids' own buffers (no copy kernel)
    if (int rc = solve_batch(b, din.data(), opt, b->d_xy.data(), nullptr, stats)) return rc;
    for (int p = 0; p < b->count; ++p) {
      OTM_HIP(hipMemcpyAsync(fwd_xy[p], b->d_xy[p], 2 * b->f[p].V * sizeof(double), hipMemcpyDeviceToHost, s));
      if (dadj[p])
        OTM_HIP(hipMemcpyAsync(adjusted_density[p], b->f[p].rho, b->f[p].F * sizeof(double), hipMemcpyDeviceToHost,
                               s));
    }
    OTM_HIP(hipStreamSynchronize(s));
    return OTM_OK;
  } catch (const HipError& e) {
    return fail_hip(e);
  }
}

int otm_batch_trace(const otm_batch* b, int job, double* residual, double* alpha, double* beta, int cap) {
  if (!b || job < 0 || job >= b->count) { set_error("invalid batch job"); return OTM_ERR_ARG; }
  try {
    OTM_HIP(hipSetDevice(b->device));
    const Fields& f = b->f[job];
    SolveState hs;
    OTM_HIP(hipMemcpy(&hs, f.state, sizeof(SolveState), hipMemcpyDeviceToHost));
    const int k = std::min(std::min(cap, hs.it), f.trace_cap);
    if (k <= 0) return 0;
    const int tc = f.trace_cap;
    if (residual) OTM_HIP(hipMemcpy(residual, f.trace, k * sizeof(double), hipMemcpyDeviceToHost));
    if (alpha) OTM_HIP(hipMemcpy(alpha, f.trace + tc, k * sizeof(double), hipMemcpyDeviceToHost));
    if (beta) OTM_HIP(hipMemcpy(beta, f.trace + 2 * tc, k * sizeof(double), hipMemcpyDeviceToHost));
    return k;
  } catch (const HipError& e) {
    return fail_hip(e);
  }
}

}  // extern "C"
