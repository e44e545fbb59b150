/*
 * gdpoisson -- MI355X (gfx950) multigrid Poisson integration, C ABI.
 *
 * Every entry point takes plain device pointers (row-major, dense, FP64 unless
 * stated) and a HIP stream passed as `void*` (NULL = the null stream).  Calls
 * enqueue work on that stream and return without waiting, except where a comment
 * says "synchronises".  Every function returns a status (GDP_OK == 0); on failure
 * gdp_last_error() gives a message for the calling thread.
 *
 * Reference interfaces replaced (paths relative to the gradient-domain package of
 * Giankaqui/visual-computing, gradient-domain/src/gradient_domain/):
 *   gdp_laplacian        operators.py:86-107   laplacian(u, spacing)         (+ Neumann variant)
 *   gdp_gradient         operators.py:30-50    gradient(image)
 *   gdp_divergence       operators.py:53-83    divergence(gx, gy)
 *   gdp_fold_boundary    operators.py:139-169  fold_boundary(rhs, boundary, spacing)
 *   gdp_smooth           multigrid.py:147-166  _smooth(u, b, spacing, sweeps)
 *   gdp_restrict         multigrid.py:169-174  _restrict(residual)
 *   gdp_prolong          multigrid.py:177-181  _prolong(correction, shape)
 *   gdp_plan_*           multigrid.py:143-144,206 level hierarchy (_coarse_shape + coarsest rule)
 *   gdp_vcycle           multigrid.py:184-215  v_cycle(u, b, spacing, options)
 *   gdp_mg_solve         multigrid.py:218-255  solve(b, spacing, options, initial)
 *                        solvers.py:121-131    _solve_multigrid
 *   gdp_mgcg_solve       solvers.py:134-160    _solve_mgcg  (SOLVERS["mgcg"])
 *   gdp_cg_solve         solvers.py:93-118     _solve_cg    (SOLVERS["cg"])
 *   gdp_relative_residual solvers.py:78-82     _relative_residual(u, b, spacing)
 *   gdp_report           solvers.py:49-75      SolverReport fields
 *   gdp_clone_rhs        poisson.py:52-92, 102-147, 261-348  seamless_clone, rectangle route:
 *                        GuidanceField.of / select / mix_by_magnitude + divergence +
 *                        fold_boundary of the window, fused
 *   gdp_guidance         poisson.py:52-92      GuidanceField.select / mix_by_magnitude / scaled
 *   gdp_illumination_field poisson.py:433-437  illumination_change's attenuated field
 *   gdp_clone_rhs_masked, gdp_masked_rhs, gdp_plan_set_mask
 *                        poisson.py:154-244    solve_masked: the system over the mask's pixels
 *                                              (fixed neighbours folded in), solved by masked
 *                                              MG-PCG instead of SuperLU
 *   gdp_tm_log_luminance hdr.py:165-167        tone_map: log of the clamped luminance
 *   gdp_tm_downsample    hdr.py:73-80          _downsample(image)
 *   gdp_tm_level_factor  hdr.py:83-104,126-133 attenuation_field: one pyramid level
 *                                              (_central_gradient, magnitude mean, local
 *                                              factor, _upsample of the coarser factor)
 *   gdp_tm_rhs           hdr.py:174-175        divergence(gx * factor, gy * factor)
 *   gdp_tm_compose       hdr.py:180-182        exp(c - max c) * ratios ** saturation
 *   gdp_percentile       hdr.py:184            numpy.percentile(image, q), linear method
 *   gdp_tm_normalise     hdr.py:185            clip(image / max(peak, 1e-12), 0, 1)
 *   gdp_neumann_dct_solve solvers.py:224-262   solve_neumann(b): idctn(dctn(b)/lambda), exact
 *   gdp_dirichlet_dst_solve solvers.py:85-90   _solve_direct (SOLVERS["direct"]): the exact
 *                        solve of sparse_laplacian(b.shape, spacing) u = b
 *                        (operators.py:110-136), by DST-I diagonalisation instead of SuperLU
 */
#ifndef GDPOISSON_H
#define GDPOISSON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDP_OK 0
#define GDP_EINVAL 1   /* bad argument: maps to KeyError / ValueError in Python   */
#define GDP_ENOMEM 2   /* workspace too small / HBM budget exceeded                */
#define GDP_EHIP 3     /* a HIP runtime call failed                                */
#define GDP_ENOTSUP 4  /* shape outside what the kernels support                   */

#define GDP_DIRICHLET 0 /* unknowns = interior, zero outside (operators.py:86-107)   */
#define GDP_NEUMANN 1   /* unknowns = all pixels, no flux across the border           */

#define GDP_FP64 0      /* V-cycle in FP64 (plain FP64 MG-PCG)                        */
#define GDP_MIXED 1     /* V-cycle in FP32; r, x, d, q, dots, coarsest solve in FP64  */

/* gdp_problem.flags */
#define GDP_PROBLEM_MASKED 1 /* a masked domain will be set (gdp_plan_set_mask): every level
                                but the coarsest is smoothed with its mask (the coarse tail
                                holds the coarsest level only), which cuts the PCG
                                iterations of masked solves by more than half */

/* One independent grid (one colour channel of one image). */
typedef struct gdp_problem {
  int32_t rows;      /* m: unknowns along axis 0 */
  int32_t cols;      /* n: unknowns along axis 1 */
  int32_t boundary;  /* GDP_DIRICHLET or GDP_NEUMANN */
  int32_t flags;     /* 0 or GDP_PROBLEM_MASKED */
  double spacing;    /* h */
} gdp_problem;

/* MultigridOptions (multigrid.py:46-71) plus the solver knobs of this library. */
typedef struct gdp_options {
  int32_t precision;        /* GDP_FP64 / GDP_MIXED                                   */
  int32_t pre_smoothing;    /* nu1 (mgcg: 1, multigrid: 2)                            */
  int32_t post_smoothing;   /* nu2                                                    */
  int32_t coarsest_size;    /* Dirichlet: min side <= this is coarsest (8);
                               Neumann:   max side <= this is coarsest (16)           */
  int32_t coarsest_sweeps;  /* Dirichlet coarsest RB-GS sweeps (60)                   */
  int32_t max_iterations;   /* PCG iterations / V-cycles                              */
  int32_t flexible;         /* 0: Fletcher-Reeves beta (reference), 1: Polak-Ribiere   */
  int32_t use_graph;        /* 1: replay PCG iterations from a captured hipGraph      */
  double tolerance;         /* relative residual target                               */
} gdp_options;

/* SolverReport (solvers.py:49-75) per problem. */
typedef struct gdp_report {
  int32_t iterations;          /* PCG iterations or V-cycles                            */
  int32_t converged;           /* 1 if the recursive residual met the target            */
  double relative_residual;    /* true ||b - A u|| / ||b|| (FP64, recomputed at the end) */
  double recursive_residual;   /* ||r_k|| / ||b|| of the last iteration                 */
} gdp_report;

typedef struct gdp_plan gdp_plan;

const char* gdp_last_error(void);
int gdp_version(void);

/* ---------------------------------------------------------------- single grids
 * `planes` independent (m, n) planes stored back to back (plane stride m*n).   */

/* out = laplacian(u, spacing): Dirichlet zero-padded, or Neumann in-grid stencil. */
int gdp_laplacian(const double* u, double* out, int m, int n, int planes,
                  double spacing, int boundary, void* stream);
/* gx, gy = gradient(img) over (h, w) planes. */
int gdp_gradient(const double* img, double* gx, double* gy, int h, int w, int planes,
                 void* stream);
/* out = divergence(gx, gy). */
int gdp_divergence(const double* gx, const double* gy, double* out, int h, int w,
                   int planes, void* stream);
/* out (h-2, w-2) = fold_boundary(rhs (h-2, w-2), boundary (h, w), spacing); out may alias rhs. */
int gdp_fold_boundary(const double* rhs, const double* boundary, double* out, int h, int w,
                      int planes, double spacing, void* stream);
/* Folded right-hand side of seamless cloning over a window (the rectangle route of
   seamless_clone, poisson.py:261-348): source planes (planes x sh x sw) placed at
   (row, col) in the target planes (planes x h x w), source-space mask (sh x sw bytes,
   nonzero = inside), guidance mode 0 import / 1 mixed / 2 average (poisson.py:319-330),
   window [top, top + hh) x [left, left + ww).  out: planes x (hh - 2) x (ww - 2) =
   divergence of the guidance at the window interior minus its ring of target values. */
int gdp_clone_rhs(const double* source, int sh, int sw, int row, int col,
                  const uint8_t* mask, const double* target, int h, int w, int planes, int top,
                  int left, int hh, int ww, int mode, double* out, void* stream);
/* The masked system's right-hand side (seamless_clone with domain "mask", poisson.py:
   154-244, 337-338): as gdp_clone_rhs, but at the window interior points where
   `interior` (h x w bytes, zero on the image's border rows and columns) is set, the
   divergence of the guidance minus the target values of the neighbours that are not
   unknowns (up, down, left, right); zero at the other points. */
int gdp_clone_rhs_masked(const double* source, int sh, int sw, int row, int col,
                         const uint8_t* mask, const double* target, int h, int w, int planes,
                         int top, int left, int hh, int ww, int mode, const uint8_t* interior,
                         double* out, void* stream);
/* The same from an explicit guidance field (gx, gy planes x h x w) and image
   (solve_masked, poisson.py:205-216). */
int gdp_masked_rhs(const double* gx, const double* gy, const double* image,
                   const uint8_t* interior, int h, int w, int planes, int top, int left, int hh,
                   int ww, double* out, void* stream);
/* Guidance-field arithmetic (GuidanceField, poisson.py:52-92) over planes x h x w:
   mode 0 select (mask[h x w] ? a : b), 1 mix_by_magnitude (|a| >= |b| ? a : b per
   component), 2 average, 3 scaled (a * factor[h x w], or a * scale if factor is NULL). */
int gdp_guidance(const double* ax, const double* ay, const double* bx, const double* by,
                 const uint8_t* mask, const double* factor, double scale, int mode, int h, int w,
                 int planes, double* out_x, double* out_y, void* stream);
/* illumination_change's attenuated field (poisson.py:433-437) over n elements:
   f = alpha ** beta * max(sqrt(gx^2 + gy^2), epsilon) ** -beta, out = (gx f, gy f). */
int gdp_illumination_field(const double* gx, const double* gy, size_t n, double alpha,
                           double beta, double epsilon, double* out_x, double* out_y,
                           void* stream);
/* ---------------------------------------------------------------- tone mapping
 * One FP64 plane per image (h x w, row-major); radiance is h x w x 3 channels-last.
 * `work` is a device buffer of gdp_tm_work_bytes() bytes (scalars and partial sums);
 * calls that share one `work` must be ordered on one stream.                       */
int gdp_tm_work_bytes(int h, int w, size_t* bytes);
/* out = log(max(radiance . (0.2126, 0.7152, 0.0722), DBL_MIN)). */
int gdp_tm_log_luminance(const double* radiance, int h, int w, double* out, void* stream);
/* out ((h+1)/2 x (w+1)/2) = every other sample of the 5-tap binomial blur of in. */
int gdp_tm_downsample(const double* in, int h, int w, double* out, void* stream);
/* factor = (s / g) (g / s)^beta with g = max(|central gradient of level|, epsilon) at
   this spacing and s = alpha max(mean g0, epsilon) (g0 the unclamped magnitude), times
   the bilinear upsampling of `coarse` (hc x wc) when it is not NULL. */
int gdp_tm_level_factor(const double* level, int h, int w, double spacing, double alpha,
                        double beta, double epsilon, const double* coarse, int hc, int wc,
                        double* factor, double* work, void* stream);
/* rhs = divergence(gx * factor, gy * factor), (gx, gy) = gradient(log_lum); h, w >= 1. */
int gdp_tm_rhs(const double* log_lum, const double* factor, int h, int w, double* rhs,
               void* stream);
/* image (h x w x 3) = (radiance / luminance)^saturation * exp(compressed - max compressed). */
int gdp_tm_compose(const double* radiance, const double* compressed, int h, int w,
                   double saturation, double* image, double* work, void* stream);
/* *out (device) = numpy.percentile(x[0 .. n), q) (linear method), n < 2^32, by an
   eight-pass radix selection; `work` holds gdp_percentile_work_bytes() bytes. */
int gdp_percentile_work_bytes(size_t* bytes);
int gdp_percentile(const double* x, size_t n, double q, double* out, void* work, void* stream);
/* image[0 .. n) = clip(image / max(*peak, 1e-12), 0, 1), peak a device scalar. */
int gdp_tm_normalise(double* image, size_t n, const double* peak, void* stream);

/* ---------------------------------------------------------------- exact Neumann solve
 * u (m x n) = idctn(dctn(b, II, ortho) / lambda, II, ortho) with coefficient (0, 0) = 0:
 * the zero-mean solution of div(grad u) = b (solve_neumann, solvers.py:224-262).  Sides
 * of at most 4096 points with prime factors 2, 3, 5 only (GDP_ENOTSUP otherwise).
 * `tables` (gdp_dct_tables_bytes, filled once per shape by gdp_dct_tables, which
 * synchronises) and `work` (gdp_dct_work_bytes) are device buffers; u may not alias b. */
int gdp_dct_supported(int m, int n);
int gdp_dct_tables_bytes(int m, int n, size_t* bytes);
int gdp_dct_tables(int m, int n, double* tables, void* stream);
int gdp_dct_work_bytes(int m, int n, size_t* bytes);
int gdp_neumann_dct_solve(const double* b, double* u, int m, int n, const double* tables,
                          double* work, void* stream);
/* The same solve without a work argument and without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under hipGraph capture,
 * serving loops that forbid hipMalloc).  The work lives in a static device array of
 * 64 MiB (2 m n doubles: m n <= 4,194,304, GDP_ENOMEM beyond), one per GPU, held until the
 * process ends; calls are identical to gdp_neumann_dct_solve's in every bit.  All calls on
 * one GPU share that array: calls that overlap on one GPU (different streams or threads
 * without ordering between them) race and give wrong results. */
int gdp_neumann_dct_solve_static(const double* b, double* u, int m, int n,
                                 const double* tables, void* stream);

/* ---------------------------------------------------------------- exact Dirichlet solve
 * u (m x n) = the solution of laplacian(u, spacing) = b with zeros outside the grid (the
 * reference's direct method, which factorises sparse_laplacian with SuperLU), by the
 * orthonormal DST-I: u = S_m ((S_m b S_n) ./ lambda) S_n.  Sides m with 2 (m + 1) <= 4096
 * and 2 (m + 1) a product of 2, 3 and 5 only (GDP_ENOTSUP otherwise).  `tables`
 * (gdp_dst_tables_bytes, filled once per shape by gdp_dst_tables, which synchronises) and
 * `work` (gdp_dct_work_bytes: 2 m n doubles) are device buffers; u may not alias b. */
int gdp_dst_supported(int m, int n);
int gdp_dst_tables_bytes(int m, int n, size_t* bytes);
int gdp_dst_tables(int m, int n, double* tables, void* stream);
int gdp_dirichlet_dst_solve(const double* b, double* u, int m, int n, double spacing,
                            const double* tables, double* work, void* stream);

/* In-place red-black Gauss-Seidel: `sweeps` sweeps, each colour 0 then 1
   (order = 0) or colour 1 then 0 (order = 1). */
int gdp_smooth(double* u, const double* b, int m, int n, double spacing, int sweeps,
               int boundary, int order, void* stream);
/* coarse (mc, nc) = 0.25 P_r^T fine P_c with fine (m, n). */
int gdp_restrict(const double* fine, double* coarse, int m, int n, int mc, int nc,
                 int boundary, void* stream);
/* fine (m, n) = P_r coarse P_c^T with coarse (mc, nc). */
int gdp_prolong(const double* coarse, double* fine, int m, int n, int mc, int nc,
                int boundary, void* stream);
/* result[p] = ||b - A u|| / ||b|| per plane (0 when b == 0).  Synchronises. */
int gdp_relative_residual(const double* u, const double* b, int m, int n, int planes,
                          double spacing, int boundary, double* result, void* stream);

/* ---------------------------------------------------------------- plans
 * A plan is a batch of independent problems solved together: every kernel launch
 * covers all of them.  The caller owns the device workspace (gdp_plan_workspace_bytes),
 * which holds the level pyramid, the FP64 Krylov vectors and the tile tables.        */
int gdp_plan_create(const gdp_problem* problems, int count, const gdp_options* options,
                    gdp_plan** plan);
int gdp_plan_workspace_bytes(const gdp_plan* plan, size_t* bytes);
/* Lays the workspace out and uploads the tables.  Synchronises. */
int gdp_plan_bind(gdp_plan* plan, void* workspace, size_t bytes);
int gdp_plan_destroy(gdp_plan* plan);
/* Number of levels of problem i and the shape of level l. */
int gdp_plan_levels(const gdp_plan* plan, int problem, int* levels);
int gdp_plan_level_shape(const gdp_plan* plan, int problem, int level, int* m, int* n);
/* Tile shapes: kind 0 = down (coarse tile), 1 = up (fine tile), 2 = Krylov (fine tile).
   Must be called before gdp_plan_bind; index into the compiled menu (see gdp_tile_menu).
   Down entries 10 .. 15 are finest-level only (GDP_EINVAL on a coarser level): the
   streamed (10 .. 12) and block-tiled (13 .. 15) down kernels that also carry the PCG
   update of an iteration in place of a separate update kernel; a plan with a masked
   domain, or with a problem in the coarse tail from its finest level, runs them as the
   plain kernels of the same shape. */
int gdp_plan_set_tile(gdp_plan* plan, int kind, int level, int menu_index);
/* Masked domain of a Dirichlet problem (solve_masked, poisson.py:154-244): `mask` is a
   device array of rows x cols bytes, nonzero = unknown; the other points are held at zero
   (their values belong in the right-hand side, see gdp_clone_rhs_masked).  The PCG
   operator and true residual act on the unknowns only; the rectangle V-cycle, masked on
   both sides, is the preconditioner.  NULL removes the mask.  Call before gdp_plan_bind;
   the array must stay valid while the plan is used. */
int gdp_plan_set_mask(gdp_plan* plan, int problem, const uint8_t* mask);
/* Block tile (rows th x columns tw) of a menu entry for a kernel with nu sweeps (the
   down and up strips get narrower as nu grows; the Krylov tiles ignore nu). */
int gdp_tile_menu(int kind, int index, int nu, int* th, int* tw);

/* u_out[i] = v_cycle(u_in[i], b[i]) for every problem (arrays of device pointers, one
   per problem, each a dense (m, n) FP64 grid).  u_out may alias u_in. */
int gdp_vcycle(gdp_plan* plan, const double* const* u_in, const double* const* b,
               double* const* u_out, void* stream);
/* Repeated V-cycles until ||b - A u|| / ||b|| < tolerance (multigrid.solve).
   `initial` may be NULL (zeros).  residual_history: max_iterations + 1 doubles per
   problem (may be NULL).  Synchronises. */
int gdp_mg_solve(gdp_plan* plan, const double* const* b, const double* const* initial,
                 double* const* u, gdp_report* reports, double* residual_history,
                 void* stream);
/* MG-preconditioned CG on -A u = -b; Neumann problems are solved for the zero-mean
   compatible part of b and return the zero-mean solution.  A problem whose recursive
   residual meets the tolerance while its FP64 true residual does not restarts from the
   true residual, at most 4 times (per problem: a problem's iterations and result do not
   depend on the other problems of the plan); reports carry the true residual.
   Synchronises. */
int gdp_mgcg_solve(gdp_plan* plan, const double* const* b, double* const* u,
                   gdp_report* reports, void* stream);
/* Jacobi-preconditioned CG (the reference's "cg" method).  Synchronises. */
int gdp_cg_solve(gdp_plan* plan, const double* const* b, double* const* u,
                 gdp_report* reports, void* stream);

/* ---------------------------------------------------------------- kernel timing
 * No counterpart in the reference (which only wall-clocks whole solves,
 * solvers.py:211-213).  With profiling on, every launch of the plan is bracketed by
 * HIP events on its launch stream (graphs are bypassed) and aggregated per kernel
 * kind with its algorithmic HBM bytes.                                           */
int gdp_plan_profile(gdp_plan* plan, int enable);  /* also clears the aggregates */
/* Synchronises the device, then reports the number of kernel kinds seen. */
int gdp_plan_profile_count(gdp_plan* plan, int* count);
int gdp_plan_profile_entry(gdp_plan* plan, int index, char* name, int name_len, double* ms,
                           long long* launches, double* bytes);

#ifdef __cplusplus
}
#endif
#endif /* GDPOISSON_H */
