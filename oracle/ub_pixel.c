/*
 * ub_pixel.c -- per-pixel mathematics of the oracle (test infrastructure).
 *
 * Restates unblending/include/unblending/blend_mode.hpp:88-202 (blend functions),
 * unblending/src/equations.cpp:10-401 (composite, energy, constraints and their
 * derivatives) and unblending/src/unblending.cpp:67-114 (objective).
 */
#include "ub_oracle.h"

#include <math.h>
#include <string.h>

enum {
    M_NORMAL = 0, M_MULTIPLY, M_SCREEN, M_OVERLAY, M_DARKEN, M_LIGHTEN, M_COLOR_DODGE,
    M_COLOR_BURN, M_HARD_LIGHT, M_SOFT_LIGHT, M_DIFFERENCE, M_EXCLUSION, M_LINEAR_DODGE
};

/* blend_mode.hpp:88 */
static const double BLEND_EPS = 1e-05;

/* blend_mode.hpp:162-196 (crop = false, as every call on the path uses it) */
double orc_blend(double s, double d, int mode)
{
    switch (mode) {
    case M_NORMAL: return s;
    case M_MULTIPLY: return s * d;
    case M_SCREEN: return 1.0 - (1.0 - s) * (1.0 - d);
    case M_OVERLAY: return (d <= 0.5) ? 2.0 * s * d : 1.0 - 2.0 * (1.0 - s) * (1.0 - d);
    case M_DARKEN: return (s < d) ? s : d;
    case M_LIGHTEN: return (s < d) ? d : s;
    case M_COLOR_DODGE: {
        if (d < BLEND_EPS) return 0.0;
        if (1.0 - s < BLEND_EPS) return 1.0;
        const double q = d / (1.0 - s);
        return (q < 1.0) ? q : 1.0; /* std::min(1.0, q) */
    }
    case M_COLOR_BURN: {
        if (1.0 - d < BLEND_EPS) return 1.0;
        if (s < BLEND_EPS) return 0.0;
        const double q = (1.0 - d) / s;
        return 1.0 - ((q < 1.0) ? q : 1.0);
    }
    case M_HARD_LIGHT: return (s <= 0.5) ? 2.0 * s * d : 1.0 - 2.0 * (1.0 - s) * (1.0 - d);
    case M_SOFT_LIGHT: {
        if (s <= 0.5) return d - (1.0 - 2.0 * s) * d * (1.0 - d);
        const double g = (d <= 0.25) ? ((16.0 * d - 12.0) * d + 4.0) * d : sqrt(d);
        return d + (2.0 * s - 1.0) * (g - d);
    }
    case M_DIFFERENCE: return (s < d) ? d - s : s - d;
    case M_EXCLUSION: return s + d - 2.0 * s * d;
    case M_LINEAR_DODGE: return s + d;
    default: return 0.0;
    }
}

/* blend_mode.hpp:90-124 */
double orc_blend_grad_s(double s, double d, int mode)
{
    switch (mode) {
    case M_NORMAL: return 1.0;
    case M_MULTIPLY: return d;
    case M_SCREEN: return 1.0 - d;
    case M_OVERLAY: return (d <= 0.5) ? 2.0 * d : 2.0 * (1.0 - d);
    case M_DARKEN: return (s < d) ? 1.0 : 0.0;
    case M_LIGHTEN: return (s < d) ? 0.0 : 1.0;
    case M_COLOR_DODGE:
        if (d < BLEND_EPS) return 0.0;
        if (1.0 - s < BLEND_EPS) return 0.0;
        if (1.0 < d / (1.0 - s)) return 0.0;
        return d / ((1.0 - s) * (1.0 - s));
    case M_COLOR_BURN:
        if (1.0 - d < BLEND_EPS) return 0.0;
        if (s < BLEND_EPS) return 0.0;
        /* -(1 < (1-d)/s ? 0 : -(1-d)/(s*s)) */
        if (1.0 < (1.0 - d) / s) return -0.0;
        return -(-(1.0 - d) / (s * s));
    case M_HARD_LIGHT: return (s <= 0.5) ? 2.0 * d : 2.0 * (1.0 - d);
    case M_SOFT_LIGHT: {
        if (s <= 0.5) return 2.0 * d * (1.0 - d);
        const double g = (d <= 0.25) ? ((16.0 * d - 12.0) * d + 4.0) * d : sqrt(d);
        return 2.0 * (g - d);
    }
    case M_DIFFERENCE: return (s < d) ? -1.0 : 1.0;
    case M_EXCLUSION: return 1.0 - 2.0 * d;
    case M_LINEAR_DODGE: return 1.0;
    default: return 0.0;
    }
}

/* blend_mode.hpp:126-160 */
double orc_blend_grad_d(double s, double d, int mode)
{
    switch (mode) {
    case M_NORMAL: return 0.0;
    case M_MULTIPLY: return s;
    case M_SCREEN: return 1.0 - s;
    case M_OVERLAY: return (d <= 0.5) ? 2.0 * s : 2.0 * (1.0 - s);
    case M_DARKEN: return (s < d) ? 0.0 : 1.0;
    case M_LIGHTEN: return (s < d) ? 1.0 : 0.0;
    case M_COLOR_DODGE:
        if (d < BLEND_EPS) return 0.0;
        if (1.0 - s < BLEND_EPS) return 0.0;
        if (1.0 < d / (1.0 - s)) return 0.0;
        return 1.0 / (1.0 - s);
    case M_COLOR_BURN:
        if (1.0 - d < BLEND_EPS) return 0.0;
        if (s < BLEND_EPS) return 0.0;
        if (1.0 < (1.0 - d) / s) return -0.0;
        return -(-1.0 / s);
    case M_HARD_LIGHT: return (s <= 0.5) ? 2.0 * s : 2.0 * (1.0 - s);
    case M_SOFT_LIGHT: {
        if (s <= 0.5) return 2.0 * s + 2.0 * d - 4.0 * s * d;
        const double h = (d <= 0.25) ? (48.0 * d * d - 24.0 * d + 4.0) : (1.0 / (2.0 * sqrt(d)));
        return 1.0 + (2.0 * s - 1.0) * h - (2.0 * s - 1.0);
    }
    case M_DIFFERENCE: return (s < d) ? 1.0 : -1.0;
    case M_EXCLUSION: return 1.0 - 2.0 * s;
    case M_LINEAR_DODGE: return 1.0;
    default: return 0.0;
    }
}

/* equations.cpp:10-32.  Products are grouped left to right as the C++
 * expressions there are. */
void orc_composite_two(const double cs[3], const double cd[3], double as, double ad,
                       const orc_layer* layer, double out[4])
{
    const double X = (double)layer->comp_x, Y = (double)layer->comp_y, Z = (double)layer->comp_z;
    const double a = X * as * ad + Y * as * (1.0 - ad) + Z * ad * (1.0 - as);
    const double ky = Y * as * (1.0 - ad);
    const double kz = Z * ad * (1.0 - as);
    for (int c = 0; c < 3; ++c) {
        const double f = orc_blend(cs[c], cd[c], layer->blend_mode);
        const double pre = f * as * ad + ky * cs[c] + kz * cd[c];
        out[c] = (a > 1e-12) ? pre / a : pre;
    }
    out[3] = a;
}

/* equations.cpp:34-74: bottom-to-top fold. */
void orc_composite_stack(const double* x, int L, const orc_layer* layers, double out[4])
{
    const double* alphas = x;
    const double* colors = x + L;
    double acc[4] = {colors[0], colors[1], colors[2], alphas[0]};
    for (int k = 1; k < L; ++k) {
        double next[4];
        orc_composite_two(colors + 3 * k, acc, alphas[k], acc[3], &layers[k], next);
        memcpy(acc, next, sizeof acc);
    }
    memcpy(out, acc, sizeof acc);
}

/* color_model.hpp:29-31: (c - mu)^T P (c - mu), row vector times matrix first. */
double orc_distance(const double c[3], const orc_layer* layer)
{
    const double* P = layer->precision;
    const double d0 = c[0] - layer->mean[0], d1 = c[1] - layer->mean[1], d2 = c[2] - layer->mean[2];
    const double t0 = d0 * P[0] + d1 * P[3] + d2 * P[6];
    const double t1 = d0 * P[1] + d1 * P[4] + d2 * P[7];
    const double t2 = d0 * P[2] + d1 * P[5] + d2 * P[8];
    return t0 * d0 + t1 * d1 + t2 * d2;
}

/* color_model.hpp:34-37: 2 P (c - mu). */
static void distance_gradient(const double c[3], const orc_layer* layer, double g[3])
{
    const double* P = layer->precision;
    const double d0 = c[0] - layer->mean[0], d1 = c[1] - layer->mean[1], d2 = c[2] - layer->mean[2];
    for (int j = 0; j < 3; ++j)
        g[j] = (2.0 * P[3 * j + 0]) * d0 + (2.0 * P[3 * j + 1]) * d1 + (2.0 * P[3 * j + 2]) * d2;
}

/* equations.cpp:104-105: sigma (sum alpha / sum alpha^2 - 1), the sums in
 * index order. */
static double sparsity_term(const double* alphas, int L, double sigma)
{
    double sum = 0.0, squared = 0.0;
    for (int i = 0; i < L; ++i) sum += alphas[i];
    for (int i = 0; i < L; ++i) squared += alphas[i] * alphas[i];
    return sigma * ((sum / squared) - 1.0);
}

/* equations.cpp:88-111: main term, the sparsity term when it is on (the
 * reference's SPARCITY build, first pass only: unblending.cpp:17,236-240),
 * the minimum-alpha term. */
double orc_energy_ex(const double* x, int L, const orc_layer* layers, int minimum_alpha, const orc_variant* v)
{
    double e = 0.0;
    for (int i = 0; i < L; ++i) e += x[i] * orc_distance(x + L + 3 * i, &layers[i]);
    if (v && v->sparsity) e += sparsity_term(x, L, v->sigma);
    if (minimum_alpha) {
        double sum = 0.0;
        for (int i = 0; i < L; ++i) sum += x[i];
        e += 0.01 * sum;
    }
    return e;
}

double orc_energy(const double* x, int L, const orc_layer* layers, int minimum_alpha)
{
    return orc_energy_ex(x, L, layers, minimum_alpha, NULL);
}

/* equations.cpp:144-152: gray rows sqrt(3) c - |c| 1 of the gray layer's
 * colour c, |c| = sqrt(c0 c0 + c1 c1 + c2 c2). */
static void gray_rows(const double* c, double* C)
{
    const double norm = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int j = 0; j < 3; ++j) C[j] = ORC_SQRT3 * c[j] - norm * 1.0;
}

int orc_gray_count(const orc_variant* v, int L)
{
    int n = 0;
    if (v)
        for (int l = 0; l < L; ++l) n += (v->gray_mask >> l) & 1;
    return n;
}

/* equations.cpp:113-155: colour rows, alpha rows, then 3 rows per gray layer
 * in layer order (the reference's SECOND_LAYER_GRAY build makes layer 1
 * gray, unblending.cpp:168-172). */
int orc_constraints_ex(const double* x, int L, const orc_layer* layers, const double target[3],
                       const double* target_alphas, const orc_variant* v, double* C)
{
    double comp[4];
    orc_composite_stack(x, L, layers, comp);
    for (int c = 0; c < 3; ++c) C[c] = comp[c] - target[c];
    int nc;
    if (target_alphas) {
        for (int i = 0; i < L; ++i) C[3 + i] = x[i] - target_alphas[i];
        nc = 3 + L;
    } else {
        C[3] = comp[3] - 1.0;
        nc = 4;
    }
    if (v)
        for (int l = 0; l < L; ++l)
            if ((v->gray_mask >> l) & 1) {
                gray_rows(x + L + 3 * l, C + nc);
                nc += 3;
            }
    return nc;
}

int orc_constraints(const double* x, int L, const orc_layer* layers, const double target[3],
                    const double* target_alphas, double* C)
{
    return orc_constraints_ex(x, L, layers, target, target_alphas, NULL, C);
}

/* equations.cpp:157-196, including the first-pass assignment (not addition)
 * of 0.01 to the alpha block at :189-193. */
void orc_energy_gradient(const double* x, int L, const orc_layer* layers, int minimum_alpha,
                         double* grad)
{
    for (int i = 0; i < L; ++i) {
        const double* c = x + L + 3 * i;
        double g[3];
        grad[i] = orc_distance(c, &layers[i]);
        distance_gradient(c, &layers[i], g);
        for (int j = 0; j < 3; ++j) grad[L + 3 * i + j] = x[i] * g[j];
    }
    if (minimum_alpha)
        for (int i = 0; i < L; ++i) grad[i] = 0.01;
}

/* ---- the reference's Jacobian assembly (equations.cpp:198-401) ----
 * A 4x4 block M[in][out]: rows are (c_r, c_g, c_b, a) of the input layer,
 * columns (B_r, B_g, B_b, A) of the composite. */
typedef struct { double m[4][4]; } mat4;

static mat4 mat4_identity(void)
{
    mat4 r;
    memset(&r, 0, sizeof r);
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0;
    return r;
}

static mat4 mat4_mul(const mat4* a, const mat4* b)
{
    mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a->m[i][k] * b->m[k][j];
            r.m[i][j] = s;
        }
    return r;
}

/* equations.cpp:198-208 */
static double dA_das(double ad, const orc_layer* l)
{
    return (double)l->comp_x * ad + (double)l->comp_y * (1.0 - ad) - (double)l->comp_z * ad;
}
static double dA_dad(double as, const orc_layer* l)
{
    return (double)l->comp_x * as - (double)l->comp_y * as + (double)l->comp_z * (1.0 - as);
}

/* equations.cpp:222-251 */
static mat4 d_by_source(const double xs[4], const double xd[4], const orc_layer* l)
{
    double xm[4];
    orc_composite_two(xs, xd, xs[3], xd[3], l, xm);
    const double A = xm[3];
    const double Y = (double)l->comp_y, Z = (double)l->comp_z;
    const double pA = dA_das(xd[3], l);
    mat4 r;
    memset(&r, 0, sizeof r);
    for (int c = 0; c < 3; ++c) {
        const double D = orc_blend(xs[c], xd[c], l->blend_mode);
        const double gs = orc_blend_grad_s(xs[c], xd[c], l->blend_mode);
        const double pC_pcs = xs[3] * xd[3] * gs + Y * (1.0 - xd[3]) * xs[3];
        r.m[c][c] = pC_pcs / A;
        const double pC_pas = D * xd[3] + Y * (1.0 - xd[3]) * xs[c] - Z * xd[3] * xd[c];
        r.m[3][c] = (pC_pas - xm[c] * pA) / A;
    }
    r.m[3][3] = pA;
    return r;
}

/* equations.cpp:253-282 */
static mat4 d_by_destination(const double xs[4], const double xd[4], const orc_layer* l)
{
    double xm[4];
    orc_composite_two(xs, xd, xs[3], xd[3], l, xm);
    const double A = xm[3];
    const double Y = (double)l->comp_y, Z = (double)l->comp_z;
    const double pA = dA_dad(xs[3], l);
    mat4 r;
    memset(&r, 0, sizeof r);
    for (int c = 0; c < 3; ++c) {
        const double D = orc_blend(xs[c], xd[c], l->blend_mode);
        const double gd = orc_blend_grad_d(xs[c], xd[c], l->blend_mode);
        const double pC_pcd = xs[3] * xd[3] * gd + Z * (1.0 - xs[3]) * xd[3];
        r.m[c][c] = pC_pcd / A;
        const double pC_pad = D * xs[3] - Y * xs[3] * xs[c] + Z * (1.0 - xs[3]) * xd[c];
        r.m[3][c] = (pC_pad - xm[c] * pA) / A;
    }
    r.m[3][3] = pA;
    return r;
}

/* equations.cpp:288-333: d xhat_k / d x_i by recursion over k. */
static mat4 d_k_by_i(const double* x, int L, const orc_layer* layers, int i, int k)
{
    if (i == 0 && k == 0) return mat4_identity();
    const double xk[4] = {x[L + 3 * k], x[L + 3 * k + 1], x[L + 3 * k + 2], x[k]};
    double prefix[4];
    {
        /* composite of layers 0..k-1 */
        double acc[4] = {x[L], x[L + 1], x[L + 2], x[0]};
        for (int j = 1; j < k; ++j) {
            double next[4];
            orc_composite_two(x + L + 3 * j, acc, x[j], acc[3], &layers[j], next);
            memcpy(acc, next, sizeof acc);
        }
        memcpy(prefix, acc, sizeof acc);
    }
    if (i == k) return d_by_source(xk, prefix, &layers[k]);
    const mat4 d1 = d_k_by_i(x, L, layers, i, k - 1);
    const mat4 d2 = d_by_destination(xk, prefix, &layers[k]);
    return mat4_mul(&d1, &d2);
}

/* equations.cpp:385-398: the gray rows' block sqrt(3) I - (1 / |c|) [c c c]
 * (entry (i, j) = sqrt(3) [i == j] - c_i / |c|), zero when |c| <= 1e-3. */
static void gray_block(const double* c, double B[3][3])
{
    const double norm = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (!(norm > 1e-3)) {
        memset(B, 0, sizeof(double) * 9);
        return;
    }
    const double inv = 1.0 / norm;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i][j] = ORC_SQRT3 * (i == j ? 1.0 : 0.0) - inv * c[i];
}

int orc_constraint_jacobian_ex(const double* x, int L, const orc_layer* layers, int refine, const orc_variant* v,
                               double* J)
{
    const int base = refine ? 3 + L : 4;
    const int nc = base + 3 * orc_gray_count(v, L);
    memset(J, 0, sizeof(double) * (size_t)(4 * L) * (size_t)nc);
    for (int i = 0; i < L; ++i) {
        const mat4 d = d_k_by_i(x, L, layers, i, L - 1);
        if (refine) {
            for (int c = 0; c < 3; ++c) J[i * nc + c] = d.m[3][c];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) J[(L + 3 * i + r) * nc + c] = d.m[r][c];
            J[i * nc + 3 + i] = 1.0;
        } else {
            for (int c = 0; c < 4; ++c) J[i * nc + c] = d.m[3][c];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 4; ++c) J[(L + 3 * i + r) * nc + c] = d.m[r][c];
        }
    }
    int q = base;
    if (v)
        for (int l = 0; l < L; ++l)
            if ((v->gray_mask >> l) & 1) {
                double B[3][3];
                gray_block(x + L + 3 * l, B);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) J[(L + 3 * l + r) * nc + q + c] = B[r][c];
                q += 3;
            }
    return nc;
}

/* equations.cpp:335-401 */
int orc_constraint_jacobian(const double* x, int L, const orc_layer* layers, int refine, double* J)
{
    const int nc = refine ? 3 + L : 4;
    memset(J, 0, sizeof(double) * (size_t)(4 * L) * (size_t)nc);
    for (int i = 0; i < L; ++i) {
        const mat4 d = d_k_by_i(x, L, layers, i, L - 1);
        if (refine) {
            for (int c = 0; c < 3; ++c) J[i * nc + c] = d.m[3][c];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) J[(L + 3 * i + r) * nc + c] = d.m[r][c];
            J[i * nc + 3 + i] = 1.0;
        } else {
            for (int c = 0; c < 4; ++c) J[i * nc + c] = d.m[3][c];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 4; ++c) J[(L + 3 * i + r) * nc + c] = d.m[r][c];
        }
    }
    return nc;
}

/* unblending.cpp:67-114.  f = E + (-lambda.C) + 0.5 rho |C|^2 and
 * grad = dE + J (rho C - lambda).  J (rho C - lambda) is accumulated by a
 * reverse sweep through the same 4x4 blocks as orc_constraint_jacobian, with
 * the block entries' common divisor A applied as one reciprocal per layer.
 * The energy terms are computed during the forward sweep. */
double orc_objective(const double* x, int L, const orc_layer* layers, const double target[3],
                     const double* target_alphas, const double* lambda, double rho,
                     double* grad, double* C)
{
    return orc_objective_ex(x, L, layers, target, target_alphas, lambda, rho, NULL, grad, C);
}

/* With the variant's terms: the sparsity term in the energy (the reference
 * turns it on in the first pass only, unblending.cpp:236-240, where the
 * alpha block's overwrite with 0.01 discards its gradient,
 * equations.cpp:177-193), and the gray rows in C with their part of
 * J (rho C - lambda), sum_j B[i][j] w_j in j order, added to the gray
 * colour's gradient after the sweep. */
double orc_objective_ex(const double* x, int L, const orc_layer* layers, const double target[3],
                        const double* target_alphas, const double* lambda, double rho, const orc_variant* v,
                        double* grad, double* C)
{
    const int refine = target_alphas != NULL;
    const double* alphas = x;
    const double* colors = x + L;

    /* forward composite keeping every prefix (equations.cpp:56-71), with the
     * energy (equations.cpp:88-111) and its colour gradient (:168-175) */
    double pre[ORC_MAX_LAYERS][4];
    double dist[ORC_MAX_LAYERS];
    double energy = 0.0;
    for (int k = 0; k < L; ++k) {
        const double* ck = colors + 3 * k;
        if (k == 0) {
            pre[0][0] = ck[0];
            pre[0][1] = ck[1];
            pre[0][2] = ck[2];
            pre[0][3] = alphas[0];
        } else {
            orc_composite_two(ck, pre[k - 1], alphas[k], pre[k - 1][3], &layers[k], pre[k]);
        }
        dist[k] = orc_distance(ck, &layers[k]);
        energy += alphas[k] * dist[k];
        double g[3];
        distance_gradient(ck, &layers[k], g);
        if (grad)
            for (int j = 0; j < 3; ++j) grad[L + 3 * k + j] = alphas[k] * g[j];
    }

    /* constraints (equations.cpp:113-155) */
    int nc;
    for (int c = 0; c < 3; ++c) C[c] = pre[L - 1][c] - target[c];
    if (refine) {
        for (int i = 0; i < L; ++i) C[3 + i] = alphas[i] - target_alphas[i];
        nc = 3 + L;
    } else {
        C[3] = pre[L - 1][3] - 1.0;
        nc = 4;
    }
    const int nbase = nc;
    if (v)
        for (int l = 0; l < L; ++l)
            if ((v->gray_mask >> l) & 1) {
                gray_rows(colors + 3 * l, C + nc);
                nc += 3;
            }
    if (v && v->sparsity) energy += sparsity_term(alphas, L, v->sigma);
    if (!refine) {
        double sum = 0.0;
        for (int i = 0; i < L; ++i) sum += alphas[i];
        energy += 0.01 * sum;
    }
    double lagrange = 0.0, squared = 0.0;
    for (int j = 0; j < nc; ++j) lagrange += lambda[j] * C[j];
    for (int j = 0; j < nc; ++j) squared += C[j] * C[j];
    const double f = energy + (-lagrange) + 0.5 * rho * squared;

    if (!grad) return f;

    double w[ORC_MAX_CONSTRAINTS];
    for (int j = 0; j < nc; ++j) w[j] = rho * C[j] - lambda[j];

    /* reverse sweep: adjoint of the composite (B, A).  The alpha block of dE
     * is D(c_i) in refinement and the reference's constant 0.01 in the first
     * pass (equations.cpp:189-193). */
    double aB[3] = {w[0], w[1], w[2]};
    double aA = refine ? 0.0 : w[3];
    for (int k = L - 1; k >= 1; --k) {
        const orc_layer* l = &layers[k];
        const double* cs = colors + 3 * k;
        const double* cd = pre[k - 1];
        const double as = alphas[k], ad = pre[k - 1][3];
        const double invA = 1.0 / pre[k][3];
        const double Y = (double)l->comp_y, Z = (double)l->comp_z;
        const double pA_s = dA_das(ad, l);
        const double pA_d = dA_dad(as, l);
        double adj_as = 0.0, adj_ad = 0.0;
        double nB[3];
        for (int c = 0; c < 3; ++c) {
            const double D = orc_blend(cs[c], cd[c], l->blend_mode);
            const double gs = orc_blend_grad_s(cs[c], cd[c], l->blend_mode);
            const double gd = orc_blend_grad_d(cs[c], cd[c], l->blend_mode);
            const double B = pre[k][c];
            const double dB_dcs = (as * ad * gs + Y * (1.0 - ad) * as) * invA;
            const double dB_das = ((D * ad + Y * (1.0 - ad) * cs[c] - Z * ad * cd[c]) - B * pA_s) * invA;
            const double dB_dcd = (as * ad * gd + Z * (1.0 - as) * ad) * invA;
            const double dB_dad = ((D * as - Y * as * cs[c] + Z * (1.0 - as) * cd[c]) - B * pA_d) * invA;
            grad[L + 3 * k + c] = grad[L + 3 * k + c] + dB_dcs * aB[c];
            adj_as += dB_das * aB[c];
            adj_ad += dB_dad * aB[c];
            nB[c] = dB_dcd * aB[c];
        }
        double ga = adj_as + pA_s * aA;
        if (refine) ga = ga + w[3 + k];
        grad[k] = (refine ? dist[k] : 0.01) + ga;
        aA = adj_ad + pA_d * aA;
        aB[0] = nB[0];
        aB[1] = nB[1];
        aB[2] = nB[2];
    }
    double ga0 = aA;
    if (refine) ga0 = ga0 + w[3];
    grad[0] = (refine ? dist[0] : 0.01) + ga0;
    for (int c = 0; c < 3; ++c) grad[L + c] = grad[L + c] + aB[c];
    int q = nbase;
    if (v)
        for (int l = 0; l < L; ++l)
            if ((v->gray_mask >> l) & 1) {
                double B[3][3];
                gray_block(colors + 3 * l, B);
                for (int i = 0; i < 3; ++i) {
                    double t = 0.0;
                    for (int j = 0; j < 3; ++j) t = t + B[i][j] * w[q + j];
                    grad[L + 3 * l + i] = grad[L + 3 * l + i] + t;
                }
                q += 3;
            }
    return f;
}
