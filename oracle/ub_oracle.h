/*
 * ub_oracle.h -- CPU FP64 restatement of Unblending's two per-pixel passes.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in unblending_amd/ is compared against, and the CPU
 * baseline that bench.py times beside it.  Nothing in the product path
 * (unblending_amd/, include/) may include, link or call it.
 *
 * What it restates (file:line into the reference TheDucker1/unblending):
 *   blend / blend_grad_s / blend_grad_d ....... unblending/include/unblending/blend_mode.hpp:88-202
 *   two-layer and stack composite ............. unblending/src/equations.cpp:10-74
 *   Gaussian colour-model distance, gradient .. unblending/include/unblending/color_model.hpp:29-37
 *   energy / constraints / derivatives ........ unblending/src/equations.cpp:76-401
 *   objective f = E - l.C + rho/2 |C|^2 ........ unblending/src/unblending.cpp:67-114
 *   default initial solution .................. unblending/src/unblending.cpp:116-153
 *   augmented-Lagrangian outer loop ........... unblending/src/unblending.cpp:155-309
 *   alpha normalisation ....................... unblending/src/unblending.cpp:311-333
 *   refinement pass structure ................. unblending/src/unblending.cpp:335-439
 *   first pass structure ...................... unblending/src/unblending.cpp:441-475
 *   clamp-edge convolution / box filter ....... unblending/src/image_processing.cpp:120-151,
 *                                               unblending/include/unblending/image_processing.hpp:215-222
 *   colour guided filter ...................... unblending/src/image_processing.cpp:214-273
 *   image-level recomposite ................... unblending/src/unblending.cpp:638-664
 *
 * The inner minimiser.  The reference calls NLopt LD_LBFGS (Luksan's PLIS)
 * through nlopt-util (unblending.cpp:249, maxeval 1000, ftol_rel = xtol_rel =
 * 5e-3).  NLopt is an un-vendored submodule (.gitmodules:7-9, pinned commit not
 * recoverable from the mounted copy) and cannot be built here, so its iterates
 * are NOT reproduced: the oracle uses the documented projected L-BFGS of
 * ub_solve.c (bounded, monotone Armijo backtracking along the projected path,
 * NLopt's relative ftol/xtol tests and evaluation cap).  Per-pixel parity
 * against NLopt is therefore "parity unpinned"; what pins this minimiser are
 * its invariants (bounds, monotone decrease, stopping rules), tested in
 * tests/test_oracle_solver.py.  Everything else above is pinned by closed
 * forms, finite differences, brute force and invariants in tests/.
 *
 * Floating point: built with -ffp-contract=off; every sum is evaluated in the
 * order written here so that a device implementation that evaluates the same
 * expressions in the same order reproduces the results to the last bit.
 */
#ifndef UB_ORACLE_H
#define UB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LAYERS 16
#define ORC_MAX_VARS (4 * ORC_MAX_LAYERS)
#define ORC_MAX_CONSTRAINTS (3 + 4 * ORC_MAX_LAYERS) /* colour, alpha and gray rows */
#define ORC_SQRT3 1.7320508075688772                   /* std::sqrt(3), correctly rounded */

/* Same field meaning as unblending_layer (unblending/include/unblending.h:67-72). */
typedef struct orc_layer {
    int32_t comp_x, comp_y, comp_z;
    int32_t blend_mode;
    double mean[3];
    double precision[9]; /* row-major */
} orc_layer;

/* Solver settings.  Defaults (orc_default_settings) are the reference's
 * constants, unblending.cpp:220-249, plus the inner minimiser's own. */
typedef struct orc_settings {
    int32_t memory;        /* L-BFGS pairs kept; <= 0: NLopt's default vector
                              storage (orc_lbfgs_memory), the default */
    int32_t max_evals;     /* per inner solve (reference: 1000) */
    int32_t max_count;     /* outer loop stops when count > max_count (reference: 20) */
    int32_t max_ls;        /* line-search trials per iteration */
    double ftol_rel;       /* 5e-3 */
    double xtol_rel;       /* 5e-3 */
    double epsilon;        /* outer tolerance, 5e-3 */
    double gamma;          /* 0.25 */
    double beta;           /* 10 */
    double initial_rho;    /* 100 */
    int32_t alpha_fixed;   /* refinement: 0 = alphas held by equality constraints
                              (reference CPU), 1 = alphas fixed by bounds
                              (reference MATTE_ALPHA_FORCE_FIXED, unblending.cpp:177-183) */
    int32_t gray_mask;     /* bit l: layer l's colour constrained to gray (the reference's
                              SECOND_LAYER_GRAY build: bit 1, unblending.cpp:168-172); 0 */
    int32_t sparsity;      /* 1: the sparsity term in the first pass's energy (the reference's
                              SPARCITY build, unblending.cpp:236-240); 0 */
    double sigma;          /* sparsity weight (unblending.cpp:234: 10) */
} orc_settings;

/* The problem variants of the reference's build flags, as the pixel math
 * takes them (sparsity is turned off by the solver in refinement). */
typedef struct orc_variant {
    int32_t gray_mask;
    int32_t sparsity;
    double sigma;
} orc_variant;

/* Per-pixel work counters (summed by the pass drivers). */
typedef struct orc_counters {
    int64_t evals;        /* value+gradient evaluations */
    int64_t iterations;   /* search directions computed (two-loop recursions) */
    int64_t outer;        /* outer AL iterations (inner solves) */
    int64_t pixels;
    int64_t max_pairs;    /* most (s, y) pairs one inner solve held */
    int64_t pairs_used;   /* (s, y) pairs the two-loop recursions went through */
} orc_counters;

void orc_default_settings(orc_settings* s);
/* Pairs the inner minimiser keeps for n variables: memory when > 0, else
 * NLopt's default for LD_LBFGS, max(1310720 / n, 10) lowered to max_evals
 * (1000 at the reference's maxeval: the whole history of an inner solve). */
int orc_lbfgs_memory(int n, int max_evals, int memory);
/* The inner minimiser's dot product for its two-loop recursion: partial sums
 * p_j of components [w j, w j + w) (w = 4 for n <= 32, else ceil(n / 8)),
 * each summed in index order from 0.0, then
 * ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)). */
double orc_tree_dot(const double* a, const double* b, int n);
/* The two-loop recursion alone: d = H q for count (s, y) pairs given oldest
 * first (S, Y: count x n, row-major), rho_k = 1 / (s_k . y_k) in index order,
 * H0 = gamma I.  Exposed to pin the recursion. */
void orc_lbfgs_direction(const double* S, const double* Y, int count, double gamma, const double* q, int n,
                         double* d);

/* ---- pixel math ---- */
double orc_blend(double s, double d, int mode);
double orc_blend_grad_s(double s, double d, int mode);
double orc_blend_grad_d(double s, double d, int mode);
/* Two-layer composite: out = (B_r, B_g, B_b, A). */
void orc_composite_two(const double cs[3], const double cd[3], double as, double ad,
                       const orc_layer* layer, double out[4]);
/* x = [alpha_0..alpha_{L-1}, c_0 (rgb) .. c_{L-1} (rgb)]. */
void orc_composite_stack(const double* x, int L, const orc_layer* layers, double out[4]);
double orc_distance(const double c[3], const orc_layer* layer);
double orc_energy(const double* x, int L, const orc_layer* layers, int minimum_alpha);
/* Returns the number of constraints (4 first pass, 3 + L refinement). */
int orc_constraints(const double* x, int L, const orc_layer* layers, const double target[3],
                    const double* target_alphas, double* C);
/* dE/dx with the reference's first-pass overwrite of the alpha block
 * (equations.cpp:189-193) when minimum_alpha != 0. */
void orc_energy_gradient(const double* x, int L, const orc_layer* layers, int minimum_alpha,
                         double* grad);
/* Transposed constraint Jacobian J[v * nc + j] = dC_j / dx_v, assembled by the
 * reference's recursion (equations.cpp:198-401).  Used to pin the adjoint. */
int orc_constraint_jacobian(const double* x, int L, const orc_layer* layers, int refine,
                            double* J);
/* Augmented-Lagrangian objective and gradient.  The gradient's constraint part
 * J.(rho C - lambda) is accumulated by a reverse sweep over the composite
 * chain.  refine != 0 selects the target-alpha constraints. */
double orc_objective(const double* x, int L, const orc_layer* layers, const double target[3],
                     const double* target_alphas, const double* lambda, double rho,
                     double* grad, double* C);

/* The same with the problem variants (v may be NULL: none).  Constraints:
 * colour rows, alpha rows, then sqrt(3) c - |c| 1 for each gray layer in
 * layer order (equations.cpp:144-152); the Jacobian's gray block
 * sqrt(3) I - c 1^T / |c|, zero when |c| <= 1e-3 (:385-398); the energy's
 * sparsity term sigma (sum alpha / sum alpha^2 - 1) (:104-105). */
int orc_gray_count(const orc_variant* v, int L);
double orc_energy_ex(const double* x, int L, const orc_layer* layers, int minimum_alpha, const orc_variant* v);
int orc_constraints_ex(const double* x, int L, const orc_layer* layers, const double target[3],
                       const double* target_alphas, const orc_variant* v, double* C);
int orc_constraint_jacobian_ex(const double* x, int L, const orc_layer* layers, int refine, const orc_variant* v,
                               double* J);
double orc_objective_ex(const double* x, int L, const orc_layer* layers, const double target[3],
                        const double* target_alphas, const double* lambda, double rho, const orc_variant* v,
                        double* grad, double* C);

/* ---- per-pixel solve ---- */
/* Solves one pixel.  init: NULL for the default first-pass initialisation,
 * else the refinement initial colours (3L) with target_alphas (L).
 * bg_color: smoothed background colour (smooth refinement) or NULL.
 * x_out receives 4L values.  Returns 1 when the outer loop stopped at
 * max_count without meeting both tolerances, else 0. */
int orc_solve_pixel(const double target[3], int L, const orc_layer* layers, int opaque,
                    const double* init_colors, const double* target_alphas,
                    const double* bg_color, const orc_settings* settings, double* x_out,
                    orc_counters* counters);
/* First pass started from x_init (4L values) and, when lambda_init is not
 * NULL, from those 4 multipliers instead of the default initialisation
 * (warm start); alpha_0 = 1 when opaque.  Returns as orc_solve_pixel. */
int orc_solve_pixel_warm(const double target[3], int L, const orc_layer* layers, int opaque,
                         const double* x_init, const double* lambda_init, const orc_settings* settings,
                         double* x_out, orc_counters* counters);
/* The inner minimiser alone (exposed for its invariant tests).  Returns the
 * most (s, y) pairs the history held; *pairs_used_out (may be NULL) receives
 * the pairs its two-loop recursions went through. */
typedef double (*orc_fn)(const double* x, double* grad, void* data);
int orc_minimize(double* x, const double* lb, const double* ub, int n, orc_fn fn, void* data,
                 const orc_settings* settings, double* f_out, int32_t* evals_out,
                 int32_t* iterations_out, int64_t* pairs_used_out);

/* ---- filter ---- */
/* Box filter, clamp-edge, radius r.  brute = 1 uses the reference's
 * O((2r+1)^2) per-pixel convolution (image_processing.cpp:120-151); brute = 0
 * a separable sliding-window sum divided by (2r+1)^2. */
void orc_box(const double* in, int w, int h, int r, int brute, double* out);
/* Colour guided filter of plane p with interleaved RGB guidance (3 doubles a
 * pixel, row stride in doubles). */
void orc_guided_filter(const double* p, const double* rgb, size_t stride, int w, int h, int r,
                       double eps, int brute, double* q);
/* The device filter's definition: the same statistics from exact integer sums
 * of values quantised to per-quantity grids (see ub_filter.c).  imax / pmax:
 * the bounds max|I| and max|p| the grids derive from (<= 0: taken from the
 * inputs; a band of an image passes the whole image's). */
void orc_guided_filter_exact(const double* p, const double* rgb, size_t stride, int w, int h, int r, double eps,
                             double imax, double pmax, double* q);
int orc_qexp(double x);
int orc_qprefix(int r);
int orc_qscale(int e, int r);
/* sqrt(3) rounded up: the coefficient bound's constant */
#define ORC_SQRT3_UP 1.7320508075688774
void orc_coeff_bounds(double imax, double pmax, double eps, double* ba, double* bb);
double orc_max_abs(const double* v, size_t n, size_t step);
int orc_default_radius(int w, int h);

/* ---- passes ---- */
/* RGB: row-major triples, stride in doubles (0 = 3w).  Output: packed
 * [layer][row][col][RGBA].  threads <= 0 selects the hardware concurrency. */
int orc_unmix(const double* rgb, size_t stride, int w, int h, const orc_layer* layers, int L,
              int opaque, const orc_settings* settings, int threads, double* out,
              orc_counters* counters);
/* Filtered, cropped and normalised target alphas [L][h*w] and (smooth) the
 * cropped smoothed background [3][h*w]; radius < 0 selects the reference's.
 * filter: ORC_FILTER_SLIDING / _BRUTE (FP64 sums) or _EXACT (the device's
 * exact-sum filter); bounds (exact filter, may be NULL): max|I| then max|p|
 * of each filtered plane (L alphas, then 3 background channels). */
enum { ORC_FILTER_SLIDING = 0, ORC_FILTER_BRUTE = 1, ORC_FILTER_EXACT = 2 };
int orc_refine_targets(const double* rgb, size_t stride, int w, int h, const orc_layer* layers,
                       int L, const double* first, int opaque, int smooth, int radius,
                       double eps, int filter, const double* bounds, double* targets, double* bg);
int orc_refine(const double* rgb, size_t stride, int w, int h, const orc_layer* layers, int L,
               const double* first, int opaque, int smooth, int radius, double eps, int filter,
               const double* bounds, const orc_settings* settings, int threads, double* out,
               orc_counters* counters);
/* Recomposite packed layers into an RGBA image [h*w][4]. */
void orc_recomposite(const double* packed, int w, int h, const orc_layer* layers, int L,
                     double* rgba);

#ifdef __cplusplus
}
#endif
#endif
