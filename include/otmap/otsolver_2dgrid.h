// Drop-in for the reference's otlib/otsolver_2dgrid.h (ggael/otmap):
// namespace otmap, BetaOpt, SolverOptions, GridBasedTransportSolver with
// set_verbose_level / init / solve.  Header-only over the C ABI of
// libotmap_amd.so (include/otmap_capi.h); link with -lotmap_amd.
//
// Differences a caller can observe:
//   * with Eigen on the include path solve() takes Eigen::Ref<const
//     Eigen::VectorXd>, as the reference; without it, any container with
//     data()/size() (std::vector<double>), or (pointer, size);
//   * SolverOptions carries four extra fields, `precision`, `refine_steps`,
//     `mixed_min_n` and `transform`;
//   * errors of the HIP layer are reported on std::cerr and leave an empty
//     map (the reference has no error channel: it prints and continues).
#pragma once

#include <cfloat>
#include <iostream>
#include <limits>
#include <memory>
#include <vector>

#include "../otmap_capi.h"
#include "transport_map.h"

namespace otmap {

// otlib/otsolver_2dgrid.h:33-35
enum struct BetaOpt { Zero, ConjugateJacobian };

// otlib/otsolver_2dgrid.h:37-43 (+ arithmetic of the Laplacian solve)
enum struct Precision { FP64, Mixed };
// FP64 Laplacian solves: DCT as dense GEMMs on FP64 MFMA, or by fast transforms
enum struct Transform { GEMM, FFT };

struct SolverOptions {
  BetaOpt beta = BetaOpt::ConjugateJacobian;
  int max_iter = 1000;
  double threshold = 1e-7;
  double max_ratio = std::numeric_limits<double>::max();
  Precision precision = Precision::FP64;
  int refine_steps = 1;
  int mixed_min_n = 1024;  // mixed: grids from this size up run the BF16x3 transform
  Transform transform = Transform::FFT;  // fast DCTs where the size allows, else the GEMMs
};

// otlib/otsolver_2dgrid.h:45-58
class GridBasedTransportSolver {
 public:
  explicit GridBasedTransportSolver(int device = 0) {
    if (otm_solver_create(&h_, device) != OTM_OK) report("GridBasedTransportSolver");
  }
  ~GridBasedTransportSolver() { otm_solver_destroy(h_); }
  GridBasedTransportSolver(const GridBasedTransportSolver&) = delete;
  GridBasedTransportSolver& operator=(const GridBasedTransportSolver&) = delete;

  /** adjust amount of debug info sent to std::cout */
  void set_verbose_level(int v) { otm_solver_set_verbose_level(h_, v); }

  /** Initializes the solver for the given grid size */
  void init(int n) {
    if (otm_solver_init(h_, n) != OTM_OK) report("init");
    origin_ = std::make_shared<GridMesh>(GridMesh::regular(n));
  }

  /** solve for the given density (n*n values, face order j + i*n) */
  TransportMap solve(const double* density, size_t size, SolverOptions opt) {
    const int n = otm_solver_grid_size(h_);
    if (n <= 0 || size != size_t(n) * size_t(n)) {
      std::cerr << "GridBasedTransportSolver::solve: density size does not match init(n)\n";
      return TransportMap();
    }
    otm_options o;
    otm_options_init(&o);
    o.beta = opt.beta == BetaOpt::Zero ? OTM_BETA_ZERO : OTM_BETA_CONJUGATE_JACOBIAN;
    o.max_iter = opt.max_iter;
    o.threshold = opt.threshold;
    o.max_ratio = opt.max_ratio;
    o.precision = opt.precision == Precision::Mixed ? OTM_PRECISION_MIXED : OTM_PRECISION_FP64;
    o.refine_steps = opt.refine_steps;
    o.mixed_min_n = opt.mixed_min_n;
    o.transform = opt.transform == Transform::FFT ? OTM_TRANSFORM_FFT : OTM_TRANSFORM_GEMM;
    const size_t V = size_t(n + 1) * (n + 1);
    std::vector<double> xy(2 * V);
    auto dens = std::make_shared<DensityVector>(long(n) * n);
    if (otm_solve(h_, density, &o, xy.data(), dens->data(), &last_) != OTM_OK) {
      report("solve");
      return TransportMap();
    }
    // the origin mesh is shared with every returned map; the forward mesh is a copy
    return TransportMap(origin_, std::make_shared<GridMesh>(n, xy), dens);
  }

#if OTMAP_HAS_EIGEN
  TransportMap solve(Eigen::Ref<const Eigen::VectorXd> density, SolverOptions opt = SolverOptions()) {
    const Eigen::VectorXd d = density;  // contiguous copy of any expression
    return solve(d.data(), size_t(d.size()), opt);
  }
#else
  template <class Vector>
  TransportMap solve(const Vector& density, SolverOptions opt = SolverOptions()) {
    return solve(density.data(), size_t(density.size()), opt);
  }
#endif

  /** iterations, residuals and device time of the last solve */
  const otm_stats& last_stats() const { return last_; }

 private:
  static void report(const char* what) { std::cerr << "otmap " << what << ": " << otm_last_error() << "\n"; }
  otm_solver* h_ = nullptr;
  std::shared_ptr<GridMesh> origin_;
  otm_stats last_{};
};

}  // namespace otmap
