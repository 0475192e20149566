/*
 * otmap_capi.h -- C ABI of libotmap_amd, the MI355X (gfx950) implementation of
 * otmap's grid transport solver.
 *
 * The reference (ggael/otmap) has no FFI; its boundary is the C++ API in
 * otlib/otsolver_2dgrid.h, otlib/transport_map.h and otlib/utils/rasterizer.h.
 * Every entry point below replaces one reference call (cited per function).
 * The C++ drop-in headers under include/otmap/ and the Python package
 * otmap_amd are thin layers over these functions.
 *
 * Conventions (identical to the reference):
 *   - n x n grid; face (cell) fields hold n*n values in face order j + i*n
 *     (i = x index, j = y index); an image img(row, col) vectorised
 *     column-major gives face (i = col, j = row).
 *   - vertex fields hold (n+1)^2 values in vertex order j + i*(n+1); vertex
 *     pairs (positions, gradients) are SoA: all x, then all y.
 *   - point lists are interleaved (x0, y0, x1, y1, ...).
 *   - all arithmetic is IEEE double unless OTM_PRECISION_MIXED is requested.
 * Host pointers are caller-owned; device buffers belong to the handles.
 * Every function returns OTM_OK (0) or a negative error code and, on error,
 * sets a per-thread message readable with otm_last_error().
 * Handles are not thread-safe; use one handle per host thread.
 */
#ifndef OTMAP_CAPI_H
#define OTMAP_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OTM_OK 0
#define OTM_ERR_ARG (-1)
#define OTM_ERR_HIP (-2)
#define OTM_ERR_STATE (-3)
#define OTM_ERR_ALLOC (-4)

/* BetaOpt (otlib/otsolver_2dgrid.h:33-35) */
enum { OTM_BETA_ZERO = 0, OTM_BETA_CONJUGATE_JACOBIAN = 1 };
/* arithmetic of the Laplacian solve and stencils */
enum { OTM_PRECISION_FP64 = 0, OTM_PRECISION_MIXED = 1 };
/* how the FP64 Laplacian solves run (otm_options.transform) */
enum { OTM_TRANSFORM_GEMM = 0, OTM_TRANSFORM_FFT = 1 };
/* why a solve stopped */
enum { OTM_EXIT_NONE = 0, OTM_EXIT_CONVERGED_OR_MAX_ITER = 1, OTM_EXIT_ZERO_DIRECTION = 2 };

/* SolverOptions (otlib/otsolver_2dgrid.h:37-43) plus the arithmetic mode. */
typedef struct otm_options {
  int beta;          /* OTM_BETA_CONJUGATE_JACOBIAN (default) or OTM_BETA_ZERO */
  int max_iter;      /* default 1000 */
  double threshold;  /* default 1e-7, compared with |r|^2 / A */
  double max_ratio;  /* default DBL_MAX: cap of max/min of the target density */
  int precision;     /* OTM_PRECISION_FP64 (default, the drop-in) or OTM_PRECISION_MIXED */
  int refine_steps;  /* mixed only: FP64 refinement steps per Laplacian solve (default 1, 0..8) */
  int mixed_min_n;   /* mixed only: grids with n >= mixed_min_n (and n % 64 == 0) run the BF16x3
                        transform, smaller grids the FP64 transform (default 1024; 0 = BF16x3 for
                        every supported size) */
  int transform;     /* grids on the FP64 transform: OTM_TRANSFORM_FFT (default: the DCTs by fast
                        transforms, for n = 2^a 3^b, 8 <= n <= 2048, n % 4 == 0; other sizes use
                        the GEMMs) or OTM_TRANSFORM_GEMM (the DCT as folded dense GEMMs on FP64
                        MFMA for every size).  No counterpart in the reference, whose solve is a
                        sparse Cholesky back-substitution. */
} otm_options;

typedef struct otm_stats {
  int iterations;           /* Newton iterations performed (reference "it") */
  int exit_reason;          /* OTM_EXIT_* */
  double initial_residual;  /* |r_0|^2 / A */
  double final_residual;    /* residual returned by the last line search */
  double ms_solve;          /* device time of the solve, density in -> positions out */
} otm_stats;

typedef struct otm_solver otm_solver;
typedef struct otm_map otm_map;
typedef struct otm_batch otm_batch;

/* Fills *opt with the reference defaults (otlib/otsolver_2dgrid.h:37-43). */
void otm_options_init(otm_options* opt);

/* Text of the last error on this thread ("" if none). */
const char* otm_last_error(void);

/* GridBasedTransportSolver() (otlib/otsolver_2dgrid.cpp:30-40); binds the
 * handle to HIP device `device` and creates its stream. */
int otm_solver_create(otm_solver** out, int device);
/* ~GridBasedTransportSolver() (otlib/otsolver_2dgrid.cpp:42-45) */
int otm_solver_destroy(otm_solver* s);
/* set_verbose_level (otlib/otsolver_2dgrid.h:52) */
int otm_solver_set_verbose_level(otm_solver* s, int level);
/* init(n) (otlib/otsolver_2dgrid.cpp:75-104): grid mesh + Laplacian operator.
 * A repeated init with the same n returns immediately, as in the reference. */
int otm_solver_init(otm_solver* s, int n);
/* current grid size (0 before init) */
int otm_solver_grid_size(const otm_solver* s);

/* solve(density, opt) (otlib/otsolver_2dgrid.cpp:106-245) on host buffers.
 *   density          n*n input values (face order), not modified
 *   fwd_xy           out, 2*(n+1)^2: forward mesh vertex positions (SoA)
 *   adjusted_density out, n*n (nullable): TransportMap::density()
 *   stats            out (nullable) */
int otm_solve(otm_solver* s, const double* density, const otm_options* opt, double* fwd_xy,
              double* adjusted_density, otm_stats* stats);
/* Same on device buffers (HIP pointers on the handle's device); returns when
 * the results are complete. */
int otm_solve_device(otm_solver* s, const double* d_density, const otm_options* opt, double* d_fwd_xy,
                     double* d_adjusted_density, otm_stats* stats);
/* Per-iteration residual / alpha / beta of the last solve (the values the
 * reference prints at verbose level 5, otlib/otsolver_2dgrid.cpp:653-658).
 * Returns the number of iterations written (<= cap). */
int otm_solver_trace(const otm_solver* s, double* residual, double* alpha, double* beta, int cap);

/* Kernel timing (no reference counterpart; used by bench.py for the roofline
 * figures): runs `reps` launches of one kernel of the iteration on the
 * handle's stream, each between two HIP events, on the buffers of the last
 * solve (which it overwrites), and returns the mean ms per launch. */
enum {
  OTM_KERNEL_DCT_GEMM = 0,  /* one n x n x n FP64-MFMA GEMM stage of the DCT solve */
  OTM_KERNEL_BETA = 1,      /* conjugate-Jacobian pass + beta step */
  OTM_KERNEL_DIR_COEF = 2,  /* direction, quartic coefficients + line search */
  OTM_KERNEL_UPDATE = 3,    /* psi / r / Jd update */
  OTM_KERNEL_DCT_SOLVE = 4, /* the whole FP64 Laplacian solve (4 GEMM stages) */
  OTM_KERNEL_MIXED_GEMM = 5, /* one 3-term BF16-split GEMM stage (6 MFMA products) */
  OTM_KERNEL_MIXED_SOLVE = 6, /* the whole mixed Laplacian solve (split, 8 stages, refinement) */
  OTM_KERNEL_FFT_SOLVE = 7,   /* the whole fast-DCT Laplacian solve (3 passes) */
  OTM_KERNEL_FFT_ROWS_FWD = 8, /* fast-DCT pass 1: DCT-II of the rows, transposed write */
  OTM_KERNEL_FFT_COLS = 9,     /* fast-DCT pass 2: DCT-II, / lambda, DCT-III of the columns */
  OTM_KERNEL_FFT_ROWS_INV = 10, /* fast-DCT pass 3: DCT-III of the rows + |d_hat|^2 partials */
  OTM_KERNEL_DIR_COEF_PASS = 11 /* the line-search pass alone (direction, quartic coefficients, partials) */
};
int otm_solver_profile_kernel(otm_solver* s, int kernel, int reps, double* ms_per_launch);

/* Tile shapes of the four FP64 DCT stages (no reference counterpart).  For
 * n % 64 == 0 the Laplacian solve runs the even/odd folded transform and
 * otm_solver_init times every candidate tile shape of each stage on the
 * device, keeping the fastest (cached per device and n).  The shapes differ in
 * speed only: every shape accumulates in the same order and gives
 * bit-identical results.
 *   otm_solver_gemm_config: cfg[0..3] = chosen shape per stage, returns the
 *     number of candidate shapes (0 when the plan does not fold);
 *   otm_solver_set_gemm_config: force shape `config` on `stage` (0..3). */
int otm_solver_gemm_config(const otm_solver* s, int* cfg);
int otm_solver_set_gemm_config(otm_solver* s, int stage, int config);

/* Tile configurations of the four mixed-precision DCT stages (no reference
 * counterpart): config 0 / 1 use workgroup tiles of 64 x 32 / 64 x 64
 * outputs, config 2 / 3 tiles of 128 x 128 computed by four / eight waves
 * (valid when n % 256 == 0).  The first
 * mixed solve times each valid configuration per stage and keeps the
 * fastest; every configuration gives bit-identical results.
 *   otm_solver_mixed_config: cfg[0..3] = chosen width per stage, returns the
 *     number of candidate widths (0 before the first mixed solve or when n is
 *     not a multiple of 64);
 *   otm_solver_set_mixed_config: force width `config` on `stage` (0..3). */
int otm_solver_mixed_config(const otm_solver* s, int* cfg);
int otm_solver_set_mixed_config(otm_solver* s, int stage, int config);

/* Component entry points: the reference's protected stencils, on host
 * buffers, for parity testing and for callers that drive the iteration.
 *   compute_vertex_gradients (otlib/otsolver_2dgrid.cpp:362-423): grads 2*(n+1)^2
 *   compute_residual         (otlib/otsolver_2dgrid.cpp:449-465): r n*n, *e = |r|^2/A
 *   Laplacian solve + mean removal (otlib/otsolver_2dgrid.cpp:174-179): d_hat n*n
 *   compute_1D_problem_parameters (otlib/otsolver_2dgrid.cpp:467-539): a, b n*n
 * These use the handle's buffers and clobber the state of a previous solve's
 * intermediates (not its results already copied out). */
int otm_vertex_gradients(otm_solver* s, const double* psi, double* grads);
int otm_residual(otm_solver* s, const double* adjusted_density, const double* psi, double* r, double* e);
int otm_laplacian_solve(otm_solver* s, const double* r, double* d_hat);
/* same solve in the mixed-precision form (3-term BF16 split GEMMs + `refine_steps`
 * FP64 refinement steps); n must be a multiple of 64 */
int otm_laplacian_solve_mixed(otm_solver* s, const double* r, double* d_hat, int refine_steps);
/* same solve by fast DCTs (OTM_TRANSFORM_FFT); n must satisfy its size rule */
int otm_laplacian_solve_fft(otm_solver* s, const double* r, double* d_hat);
int otm_line_search_coefficients(otm_solver* s, const double* g0, const double* dir, double* a, double* b);
/* Step choice of solve_1D_problem (otlib/otsolver_2dgrid.cpp:563-595) on the
 * device, for `count` coefficient vectors z (5 doubles each: e(t) = z0 + z1 t
 * + ... + z4 t^4, the reference's z): alpha[k] = the real stationary point in
 * [-10, 10] of least e, else 0.5; emin[k] = e(alpha[k]).  This is the choice
 * every device line search makes (host buffers; synchronous). */
int otm_line_search_steps(int device, const double* z, int count, double* alpha, double* emin);

/* ---- TransportMap (otlib/transport_map.h:29-63) -------------------------
 * A map holds the forward mesh (2*(n+1)^2 SoA positions) and the adjusted
 * density (n*n) of one solve.  The BVH of the forward quads is built on the
 * first inverse query (or by otm_map_init_inverse), as the reference's lazy
 * init_inverse() (otlib/transport_map.cpp:36-43). */
/* TransportMap(origin, fwd, density) (otlib/transport_map.cpp:30-34) on host buffers */
int otm_map_create(otm_map** out, int device, int n, const double* fwd_xy, const double* density);
/* same from device buffers (e.g. the outputs of otm_solve_device) */
int otm_map_create_device(otm_map** out, int device, int n, const double* d_fwd_xy, const double* d_density);
/* ~TransportMap() (otlib/transport_map.cpp:45-48) */
int otm_map_destroy(otm_map* m);
/* init_inverse() (otlib/transport_map.cpp:36-43): BVH2D::build(fwd, 4, 24) */
int otm_map_init_inverse(otm_map* m);
/* number of BVH nodes (0 before the build) */
int otm_map_bvh_size(const otm_map* m);
/* The BVH as built (no reference counterpart: inspection and parity tests of
 * the device build against BVH2D::build, otlib/utils/bvh2d.cpp:29-243), in
 * the reference's node order: per node its inflated box (boxes[4k .. 4k+3] =
 * min x, min y, max x, max y), first (inner: first child, leaf: first face
 * slot), count (leaf: faces, inner: -1) and escape (next node in depth-first
 * order once the subtree is done, -1 at the end); faces: the n^2 face slots.
 * Builds the BVH if needed; arrays sized by otm_map_bvh_size. */
int otm_map_bvh_export(otm_map* m, double* boxes, int32_t* first, int32_t* count, int32_t* escape, int32_t* faces);
/* inv / inv_fast / apply_inverse_map (otlib/transport_map.cpp:50-128) for
 * `count` interleaved points.  mode 0 = inv (all overlapping faces, the
 * reference's selection), 1 = inv_fast (first hit).  out: interleaved
 * positions (NaN where no face holds the point -- the reference prints an
 * error there); face (nullable): selected face index, -1 if none; nhits
 * (nullable): number of overlapping faces found. */
int otm_map_inverse(otm_map* m, const double* pts, int64_t count, int mode, double* out, int32_t* face,
                    int32_t* nhits);
/* same on device buffers */
int otm_map_inverse_device(otm_map* m, const double* d_pts, int64_t count, int mode, double* d_out,
                           int32_t* d_face, int32_t* d_nhits);

/* Inverse images of the origin grid vertices (i/n, j/n), in vertex order,
 * SoA: the inverse map as a mesh, as the reference builds it by applying
 * apply_inverse_map to a copy of origin_mesh() (apps/barycenters.cpp:187-190).
 * mode 0 = inv, 1 = inv_fast; NaN where no face holds the point. */
int otm_map_inverse_vertices(otm_map* m, int mode, double* xy);
int otm_map_inverse_vertices_device(otm_map* m, int mode, double* d_xy);

/* ---- rasterize_image (otlib/utils/rasterizer.h:21-22, rasterizer.cpp:188-308)
 * Per-face density mode on the (n+1)^2-vertex quad mesh `mesh_xy` (SoA):
 * every pixel covered by a face gets float(1/n^2) * density(f) / |area(f)|,
 * the largest covering face index wins, uncovered pixels are 0.  density may
 * be NULL (all ones: the reference's 2-argument overload).  img is row-major
 * rows x cols: img[y*cols + x] is the reference's img(y, x). */
int otm_rasterize(int device, int n, const double* mesh_xy, const double* density, int rows, int cols, double* img);
/* same on device buffers; img_f32 != 0 writes float, else double; d_winner:
 * rows*cols int32 scratch (receives the winning face index, -1 if none). */
int otm_rasterize_device(int device, int n, const double* d_mesh_xy, const double* d_density, int rows, int cols,
                         int img_f32, void* d_img, int32_t* d_winner);

/* ---- Transport cost ----------------------------------------------------
 * transport_cost(src_mesh, dst_mesh, density_per_face, cost_per_face)
 * (otlib/transport_map.cpp:131-215) on (n+1)^2-vertex quad grids (SoA
 * positions): *total = sum(cost) / integral; cost (nullable) per face,
 * normalised by the same integral.  As in the reference, each quad is costed
 * through the triangle of its first halfedge (vertices 3, 0, 1). */
int otm_transport_cost(int device, int n, const double* src_xy, const double* dst_xy, const double* density,
                       double* cost, double* total);
int otm_transport_cost_device(int device, int n, const double* d_src_xy, const double* d_dst_xy,
                              const double* d_density, double* d_cost, double* total);
/* compute_transport_cost (otlib/otsolver_2dgrid.cpp:608-640): 2 x 2 Gauss
 * points of |displacement|^2 per face times density * A, for vertex
 * displacements disp_xy (SoA, 2*(n+1)^2); *total = sum of cost. */
int otm_grid_transport_cost(int device, int n, const double* disp_xy, const double* density, double* cost,
                            double* total);

/* ---- Pipeline helpers (device buffers; no reference C++ API: the steps of
 * the reference's barycenters and stippling apps) --------------------------
 * otm_mesh_lerp_device: out = lerp (k = 2: (1-u) m0 + u m1) or bilerp (k = 4:
 *   (1-u)(1-v) m0 + u(1-v) m1 + uv m2 + (1-u)v m3) of `count` doubles,
 *   apps/barycenters.cpp:128-140, 254-277.
 * otm_image_normalize_mean_device: out = img * (expected_mean / mean(img)),
 *   written as float (out_f32 != 0) or double; *mean (nullable) = mean(img),
 *   apps/barycenters.cpp:280-289.
 * otm_prune_points_device: keeps, in order, the interleaved points whose
 *   pixel (floor(p * n)) has non-zero density (face order), or that are not
 *   NaN; *kept = number written to d_out (room for `count` points),
 *   apps/stippling.cpp:188-196 (pixel indices are clamped to the grid). */
int otm_mesh_lerp_device(int device, long count, int k, const double* const* d_meshes, double u, double v,
                         double* d_out);
int otm_image_normalize_mean_device(int device, long count, const double* d_img, double expected_mean, int out_f32,
                                    void* d_out, double* mean);
int otm_prune_points_device(int device, int n, const double* d_density, const double* d_pts, int64_t count,
                            double* d_out, int64_t* kept);
/* generate_blue_noise_tile (apps/common/image_utils.cpp:148-173, used by
 * apps/stippling.cpp:177-181) on device: the nb-point tile d_tile
 * (interleaved, coordinates in [0,1)), N = floor(sqrt(nb)) per side,
 * repeated over ceil(n/N)^2 cells scaled by N/n; the points with both
 * coordinates <= 1 are written to d_out (room for cap points) in the
 * reference's order (cell row i, cell column j, tile point k); *count = their
 * number (also when cap is too small, which returns OTM_ERR_ARG). */
int otm_blue_noise_tile_device(int device, int n, const double* d_tile, int64_t nb, double* d_out, int64_t cap,
                               int64_t* count);
/* otm_prune_points_device without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * results; its workgroup counts live in a fixed device array of the library
 * (32 KB per GPU, held until the process ends), so count is limited to
 * 16777216 points (larger counts return OTM_ERR_ARG), and calls that overlap
 * on one GPU race.  Synchronous, like otm_prune_points_device. */
int otm_prune_points_device_noalloc(int device, int n, const double* d_density, const double* d_pts, int64_t count,
                                    double* d_out, int64_t* kept);

/* ---- Batched solves -----------------------------------------------------
 * No reference counterpart: the reference solves one density per
 * GridBasedTransportSolver (otlib/otsolver_2dgrid.cpp:106-245) and its apps
 * loop over densities (apps/barycenters.cpp:167-197).  A batch owns `count`
 * grids of sizes n[p] (multiples of 64) and solves all of them together: each
 * step of a Newton iteration is one kernel launch for the whole batch, the
 * iterations are captured in a HIP graph, and the DCT tile shapes are chosen
 * by the autotuner for the batch's sizes.  Every grid gets exactly the result
 * of a single-grid solve of its density (bit-identical); stats[p] holds
 * grid p's iterations and residuals (ms_solve: the whole batch).
 *   density / fwd_xy / adjusted_density: arrays of `count` pointers, each
 *   sized for its grid as in otm_solve (adjusted_density and its entries
 *   nullable).  Options apply to every grid, both precision modes included
 *   (a mixed batch gives each grid the result of its single mixed solve). */
int otm_batch_create(otm_batch** out, int device, int count, const int* n);
int otm_batch_destroy(otm_batch* b);
int otm_batch_size(const otm_batch* b);
int otm_batch_solve(otm_batch* b, const double* const* density, const otm_options* opt, double* const* fwd_xy,
                    double* const* adjusted_density, otm_stats* stats);
int otm_batch_solve_device(otm_batch* b, const double* const* d_density, const otm_options* opt,
                           double* const* d_fwd_xy, double* const* d_adjusted_density, otm_stats* stats);
/* per-iteration residual / alpha / beta of grid `job` in the last solve */
int otm_batch_trace(const otm_batch* b, int job, double* residual, double* alpha, double* beta, int cap);
/* Tile shapes are tuned per grid size: the jobs of each size run the shape
 * the autotuner picked for that size (one launch per shape in use per stage).
 * otm_batch_gemm_config: folded FP64 tile shape per stage of grid `job` (-1:
 * of the largest grid); returns the number of candidates. */
int otm_batch_gemm_config(const otm_batch* b, int job, int* cfg);
/* tile configuration per mixed DCT stage (see otm_solver_mixed_config) of grid
 * `job` (-1: of the largest grid), picked on the first mixed solve of the
 * batch; returns the number of candidates (0 before it) */
int otm_batch_mixed_config(const otm_batch* b, int job, int* cfg);

#ifdef __cplusplus
}
#endif

#endif /* OTMAP_CAPI_H */
