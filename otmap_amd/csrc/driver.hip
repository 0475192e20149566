// Per-grid buffers and the iteration driver shared by the single-grid solver
// handle and the batched handle.
//
// A solve of a job table is: density adjustment, field initialisation and the
// initial residual (otsolver_2dgrid.cpp:106-160), then Newton iterations until
// every job's device state says done (:160-221), then the forward positions
// (:223-227).  Iterations are captured once into two HIP graphs, one of
// ITERS_PER_GRAPH iterations and one of a single iteration, and replayed;
// after each replay of a batch one small kernel counts the jobs still
// running, and those of them still far from their stop, straight into one of
// two slots of mapped pinned host memory (no copy launch); a single job's
// scalar kernels write its own two flags there when they change (no count
// kernel).  The host reads the counts of replay k while replay k + 1 runs
// (the device never waits for the host).  Long replays run while some job is
// far from its stop;
// once every running job is near it (residual within NEAR_FACTOR of the
// threshold, or within a long replay of max_iter) the host queues single
// iterations, so at most about two iterations run gated after the last job
// stops instead of up to 2 * ITERS_PER_GRAPH - 1.  Iterations replayed after a
// job has stopped are no-ops for that job (every kernel checks its gate), so
// results do not depend on how many replays are in flight or how long they are.

#include <cfloat>
#include <cstring>

#include "otm_internal.h"

namespace otm {

namespace {

template <class T>
void dalloc(T*& p, size_t count) {
  OTM_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

long partial_slots(int n) {
  const long gemm_tiles = long((n + 63) / 64) * ((n + 63) / 64);
  return std::max<long>(std::max<long>(num_tiles(n), gemm_tiles), 3 * 1024);
}

}  // namespace

void fields_create(Fields& f, int n, int trace_cap) {
  f = Fields();
  f.n = n;
  f.F = long(n) * n;
  f.V = long(n + 1) * (n + 1);
  const size_t F = f.F, V = f.V;
  dalloc(f.rho, F); dalloc(f.psi, F); dalloc(f.psi2, F); dalloc(f.g0, 2 * V); dalloc(f.r, F); dalloc(f.jd, F);
  dalloc(f.dhat, F); dalloc(f.d[0], F); dalloc(f.d[1], F); dalloc(f.dpair, 2);
  f.pslots = partial_slots(n);
  dalloc(f.a, F); dalloc(f.b, F); dalloc(f.partials, size_t(SITE_COUNT) * SITE_VALUES * size_t(f.pslots));
  dalloc(f.tmp0, F); dalloc(f.tmp1, F); dalloc(f.state, 1);
  OTM_HIP(hipMemcpy(f.dpair, f.d, 2 * sizeof(double*), hipMemcpyHostToDevice));
  OTM_HIP(hipMemset(f.state, 0, sizeof(SolveState)));
  dalloc(f.trace, 3 * size_t(trace_cap));
  f.trace_cap = trace_cap;
}

void fields_destroy(Fields& f) {
  dfree(f.rho); dfree(f.psi); dfree(f.psi2); dfree(f.g0); dfree(f.r); dfree(f.jd); dfree(f.dhat);
  dfree(f.d[0]); dfree(f.d[1]); dfree(f.dpair); dfree(f.a); dfree(f.b); dfree(f.partials);
  dfree(f.tmp0); dfree(f.tmp1); dfree(f.state); dfree(f.trace);
  dfree(f.mrhs); dfree(f.mwa); dfree(f.mwb); dfree(f.x0);
  f = Fields();
}

void fields_enable_mixed(Fields& f) {
  if (f.x0) return;
  const size_t F = f.F;
  dalloc(f.mrhs, 3 * F);
  dalloc(f.mwa, 3 * F);
  dalloc(f.mwb, 3 * F);
  dalloc(f.x0, F);
}

bool fields_grow_trace(Fields& f, int cap) {
  if (cap <= f.trace_cap) return false;
  dfree(f.trace);
  dalloc(f.trace, 3 * size_t(cap));
  f.trace_cap = cap;
  return true;
}

void driver_create(Driver& d) {
  OTM_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
  OTM_HIP(hipEventCreate(&d.ev0));
  OTM_HIP(hipEventCreate(&d.ev1));
  OTM_HIP(hipEventCreate(&d.evc[0]));
  OTM_HIP(hipEventCreate(&d.evc[1]));
  // the count kernel writes the host's slots directly (coherent: no host-side
  // cache to flush; the event after the kernel orders the host's read)
  OTM_HIP(hipHostMalloc(reinterpret_cast<void**>(&d.h_open), 4 * sizeof(int),
                        hipHostMallocMapped | hipHostMallocCoherent));
  OTM_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&d.d_open), d.h_open, 0));
}

void driver_invalidate(Driver& d) {
  for (int g = 0; g < 2; ++g) {
    if (d.exec[g]) (void)hipGraphExecDestroy(d.exec[g]);
    if (d.graph[g]) (void)hipGraphDestroy(d.graph[g]);
    d.exec[g] = nullptr;
    d.graph[g] = nullptr;
  }
  d.key = -1;
}

void driver_destroy(Driver& d) {
  if (d.stream) (void)hipStreamSynchronize(d.stream);
  driver_invalidate(d);
  if (d.h_open) (void)hipHostFree(d.h_open);
  if (d.ev0) (void)hipEventDestroy(d.ev0);
  if (d.ev1) (void)hipEventDestroy(d.ev1);
  for (hipEvent_t e : d.evc)
    if (e) (void)hipEventDestroy(e);
  if (d.stream) (void)hipStreamDestroy(d.stream);
  d = Driver();
}

// One iteration of every job: each job's Laplacian solve on its route (FP64
// GEMMs, fast DCTs, or the mixed transform, whose jobs find their folded
// right-hand side written by the previous update pass), then the shared
// beta, line-search and update passes.
void enqueue_iteration(const JobSet& js, const DctPlan* plan, const otm_options& opt, hipStream_t s, int* flags) {
  if (js.routed_mixed + js.routed_fft < js.count) launch_dct_solve(js, plan, s);
  if (js.routed_fft > 0) launch_fft_solve(js, true, s);
  if (js.routed_mixed > 0) launch_mixed_solve(js, opt.refine_steps, true, s, false);
  launch_beta(js, s, flags);
  launch_dir_coef(js, s, flags);
  launch_update(js, s);
}

int driver_solve(Driver& d, JobSet& js, const DctPlan* plan, const otm_options& opt, float* ms) {
  hipStream_t s = d.stream;
  // the captured iteration depends on the precision mode and the tile shapes
  const long key = long(opt.refine_steps) | (js.version << 8);
  if (d.exec[0] && d.key != key) driver_invalidate(d);
  jobset_upload(js, s);
  OTM_HIP(hipEventRecord(d.ev0, s));
  // a single job reports through its own flags, a batch through the count kernel
  const bool direct = js.count == 1;
  int* flags = direct ? d.d_open : nullptr;
  launch_start(js, opt.max_ratio, opt.max_iter, opt.threshold, opt.beta == OTM_BETA_CONJUGATE_JACOBIAN ? 1 : 0, s,
               flags);
  launch_mprep(js, true, s);  // jobs on the mixed transform: right-hand side of the first iteration
  if (!d.exec[0]) {
    for (int g = 0; g < 2; ++g) {
      OTM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        for (int k = 0; k < (g == 0 ? ITERS_PER_GRAPH : 1); ++k) enqueue_iteration(js, plan, opt, s, flags);
      } catch (...) {
        hipGraph_t gr = nullptr;
        (void)hipStreamEndCapture(s, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      OTM_HIP(hipStreamEndCapture(s, &d.graph[g]));
      OTM_HIP(hipGraphInstantiate(&d.exec[g], d.graph[g], nullptr, nullptr, 0));
    }
    d.key = key;
  }
  // one replay in flight ahead of the check: replay k + 1 is queued before
  // the host waits for the counts of replay k (slot k % 2); its length follows
  // the counts of replay k - 1
  long launched = 0;
  auto replay = [&](int slot, bool single) {
    OTM_HIP(hipGraphLaunch(d.exec[single ? 1 : 0], s));
    launched += single ? 1 : ITERS_PER_GRAPH;
    if (!direct) launch_count_open(js, d.d_open + 2 * slot, s);
    OTM_HIP(hipEventRecord(d.evc[slot], s));
  };
  int slot = 0;
  bool single = false;
  replay(slot, single);
  for (;;) {
    const bool more = launched <= long(opt.max_iter);
    if (more) replay(slot ^ 1, single);
    OTM_HIP(hipEventSynchronize(d.evc[slot]));
    const volatile int* cnt = d.h_open + (direct ? 0 : 2 * slot);
    if (cnt[0] == 0) break;
    if (!more) {
      set_error("internal error: solve did not reach its stop condition");
      return OTM_ERR_STATE;
    }
    single = cnt[1] == 0;
    slot ^= 1;
  }
  launch_finish(js, s);
  OTM_HIP(hipEventRecord(d.ev1, s));
  OTM_HIP(hipStreamSynchronize(s));
  float t = 0.f;
  OTM_HIP(hipEventElapsedTime(&t, d.ev0, d.ev1));
  if (ms) *ms = t;
  return OTM_OK;
}

}  // namespace otm
