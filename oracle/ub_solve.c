/*
 * ub_solve.c -- per-pixel constrained solve of the oracle (test infrastructure).
 *
 * Outer loop: the reference's augmented-Lagrangian iteration,
 * unblending/src/unblending.cpp:155-309, with its constants (:220-224) and the
 * default initial solution (:116-153, :185-208).
 *
 * Inner loop: the reference calls NLopt LD_LBFGS (unblending.cpp:249).  NLopt
 * is not available (un-vendored submodule), so this file defines the inner
 * minimiser the whole project uses -- a projected limited-memory BFGS:
 *
 *   1. free set: i is free unless lb_i == ub_i, or x_i sits on a bound with the
 *      gradient pointing outward; q = g restricted to the free set.
 *      Stop when max|q_i| == 0.
 *   2. direction: the two-loop recursion over the stored (s, y) pairs
 *      (newest to oldest, then back), initial scaling s.y / y.y of the newest
 *      pair; d = -H q on the free set, 0 elsewhere.  If g.d is not negative the
 *      history is dropped and d = -q.  The recursion's dot products are summed
 *      in a fixed tree order (orc_tree_dot): eight partial sums of four
 *      consecutive components each, then pairwise -- the order in which the
 *      device's 8-lane groups form them.
 *   3. step: t = 1, or min(1, 1/|d|) with an empty history.  Trial points
 *      x(t) = clamp(x + t d, lb, ub); accept when f(x(t)) <= f + 1e-4 g.(x(t) - x).
 *      Otherwise t <- the minimiser of the quadratic through f, g.(x(t)-x) and
 *      f(x(t)), clamped to [0.1 t, 0.5 t] (0.5 t when g.(x(t)-x) >= 0).  At most
 *      max_ls trials; a failed search ends the inner solve at the last
 *      accepted point (the monotone minimiser's best point).
 *   4. update: s = x+ - x, y = g+ - g; the pair is kept when s.y > 2^-52 y.y
 *      (memory pairs, oldest dropped), otherwise the whole history is dropped.
 *      memory <= 0 selects NLopt's default vector storage for LD_LBFGS:
 *      M = max(MEMAVAIL / n, 10) with MEMAVAIL = 1310720 doubles, lowered to
 *      maxeval when maxeval <= M (NLopt src/algs/luksan/plis.c, luksan.h; the
 *      reference never calls set_vector_storage).  At n <= 64 and maxeval =
 *      1000 that is M = 1000: no pair of an inner solve is ever dropped,
 *      since every iteration costs at least one evaluation.
 *   5. stop (NLopt's relative tests, nlopt util/stop.c semantics):
 *      |f+ - f| < ftol_rel (|f+| + |f|)/2 or f+ == f; or for every i
 *      |x+_i - x_i| < xtol_rel (|x+_i| + |x_i|)/2 or x+_i == x_i; or the
 *      evaluation count reaches max_evals; or f or g stops being finite.
 *
 * Every reduction is a left-to-right sum in index order.
 */
#include "ub_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void orc_default_settings(orc_settings* s)
{
    s->memory = 0;         /* NLopt's default vector storage (orc_lbfgs_memory) */
    s->max_evals = 1000;   /* unblending.cpp:249 */
    s->max_count = 20;     /* unblending.cpp:246 */
    s->max_ls = 20;
    s->ftol_rel = 5e-3;    /* unblending.cpp:222 local_epsilon */
    s->xtol_rel = 5e-3;
    s->epsilon = 5e-3;     /* unblending.cpp:221 */
    s->gamma = 0.25;       /* :220 */
    s->beta = 10.0;        /* :223 */
    s->initial_rho = 100.0;/* :224 */
    s->alpha_fixed = 0;    /* MATTE_ALPHA_FORCE_FIXED off, unblending.cpp:19 */
    s->gray_mask = 0;      /* SECOND_LAYER_GRAY off, unblending.cpp:21 */
    s->sparsity = 0;       /* SPARCITY off, unblending.cpp:17 */
    s->sigma = 10.0;       /* unblending.cpp:234 */
}

static double clampd(double v, double lo, double hi)
{
    v = (v < lo) ? lo : v;
    return (v > hi) ? hi : v;
}

static double dot(const double* a, const double* b, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static int relstop(double vold, double vnew, double reltol)
{
    return fabs(vnew - vold) < reltol * (fabs(vnew) + fabs(vold)) * 0.5 || vnew == vold;
}

/* components per partial sum of orc_tree_dot */
static int tree_width(int n) { return n <= 32 ? 4 : (n + 7) / 8; }

double orc_tree_dot(const double* a, const double* b, int n)
{
    const int w = tree_width(n);
    double p[8];
    for (int j = 0; j < 8; ++j) {
        double s = 0.0;
        for (int c = 0; c < w; ++c) {
            const int i = j * w + c;
            if (i < n) s = s + a[i] * b[i];
        }
        p[j] = s;
    }
    return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

int orc_lbfgs_memory(int n, int max_evals, int memory)
{
    if (memory > 0) return memory;
    int m = 1310720 / (n > 0 ? n : 1);
    if (m < 10) m = 10;
    if (max_evals > 0 && max_evals <= m) m = max_evals;
    return m;
}

/* (s, y) history of one inner solve: a ring of M pairs, held on the stack
 * while it is short and moved to the heap when it outgrows that. */
#define ORC_STACK_PAIRS 32
typedef struct {
    double* S;  /* pair k of the ring: S + k * n */
    double* Y;
    double* rh;
    double* a;
    int cap, n;
    void* heap;
} lbfgs_history;

static int history_grow(lbfgs_history* h, int head, int count, int want)
{
    const size_t pairs = (size_t)want;
    double* mem = (double*)malloc(sizeof(double) * pairs * (2 * (size_t)h->n + 2));
    if (!mem) return 0;
    double* S = mem;
    double* Y = S + pairs * h->n;
    double* rh = Y + pairs * h->n;
    double* a = rh + pairs;
    for (int k = 0; k < count; ++k) { /* oldest first */
        const int j = (head + k) % h->cap;
        memcpy(S + (size_t)k * h->n, h->S + (size_t)j * h->n, sizeof(double) * h->n);
        memcpy(Y + (size_t)k * h->n, h->Y + (size_t)j * h->n, sizeof(double) * h->n);
        rh[k] = h->rh[j];
    }
    free(h->heap);
    h->heap = mem;
    h->S = S;
    h->Y = Y;
    h->rh = rh;
    h->a = a;
    h->cap = want;
    return 1;
}

/* d = H q for the history's count pairs (the k-th oldest in slot
 * (head + k) % cap), H the L-BFGS inverse-Hessian approximation from
 * H0 = gamma I; q is overwritten. */
static void two_loop(lbfgs_history* hs, int head, int count, double gamma, double* q, double* d, int n)
{
    for (int k = count - 1; k >= 0; --k) {
        const int j = (head + k) % hs->cap;
        const double* Sj = hs->S + (size_t)j * n;
        const double* Yj = hs->Y + (size_t)j * n;
        hs->a[k] = hs->rh[j] * orc_tree_dot(Sj, q, n);
        for (int i = 0; i < n; ++i) q[i] = q[i] - hs->a[k] * Yj[i];
    }
    for (int i = 0; i < n; ++i) d[i] = gamma * q[i];
    for (int k = 0; k < count; ++k) {
        const int j = (head + k) % hs->cap;
        const double* Sj = hs->S + (size_t)j * n;
        const double* Yj = hs->Y + (size_t)j * n;
        const double b = hs->rh[j] * orc_tree_dot(Yj, d, n);
        for (int i = 0; i < n; ++i) d[i] = d[i] + (hs->a[k] - b) * Sj[i];
    }
}

void orc_lbfgs_direction(const double* S, const double* Y, int count, double gamma, const double* q, int n,
                         double* d)
{
    double* rh = (double*)malloc(sizeof(double) * (size_t)(2 * (count > 0 ? count : 1)));
    double qq[ORC_MAX_VARS];
    if (!rh) abort();
    for (int k = 0; k < count; ++k) rh[k] = 1.0 / dot(S + (size_t)k * n, Y + (size_t)k * n, n);
    for (int i = 0; i < n; ++i) qq[i] = q[i];
    lbfgs_history hs = {(double*)S, (double*)Y, rh, rh + count, count > 0 ? count : 1, n, NULL};
    two_loop(&hs, 0, count, gamma, qq, d, n);
    free(rh);
}

int orc_minimize(double* x, const double* lb, const double* ub, int n, orc_fn fn, void* data,
                 const orc_settings* st, double* f_out, int32_t* evals_out, int32_t* iterations_out,
                 int64_t* pairs_used_out)
{
    const int M = orc_lbfgs_memory(n, st->max_evals, st->memory);
    double g[ORC_MAX_VARS], d[ORC_MAX_VARS], q[ORC_MAX_VARS], xt[ORC_MAX_VARS], gt[ORC_MAX_VARS];
    double S0[ORC_STACK_PAIRS * ORC_MAX_VARS], Y0[ORC_STACK_PAIRS * ORC_MAX_VARS];
    double rh0[ORC_STACK_PAIRS], a0[ORC_STACK_PAIRS];
    lbfgs_history hs = {S0, Y0, rh0, a0, M < ORC_STACK_PAIRS ? M : ORC_STACK_PAIRS, n, NULL};
    int freev[ORC_MAX_VARS];
    int count = 0, head = 0; /* ring: slot (head + k) % cap holds the k-th oldest pair */
    double gamma = 1.0;
    int32_t max_run = 0;
    int64_t used = 0;

    for (int i = 0; i < n; ++i) x[i] = clampd(x[i], lb[i], ub[i]);
    double f = fn(x, g, data);
    int32_t nev = 1, iters = 0;

    for (;;) {
        int finite = isfinite(f);
        for (int i = 0; i < n; ++i) finite = finite && isfinite(g[i]);
        if (!finite) break;

        double pgmax = 0.0;
        for (int i = 0; i < n; ++i) {
            freev[i] = lb[i] < ub[i] && !(x[i] <= lb[i] && g[i] > 0.0) && !(x[i] >= ub[i] && g[i] < 0.0);
            q[i] = freev[i] ? g[i] : 0.0;
            const double m = fabs(q[i]);
            pgmax = (m > pgmax) ? m : pgmax;
        }
        if (pgmax == 0.0) break;
        ++iters;
        used += count;

        two_loop(&hs, head, count, gamma, q, d, n);
        for (int i = 0; i < n; ++i) d[i] = freev[i] ? -d[i] : 0.0;
        double gd = dot(g, d, n);
        if (!(gd < 0.0)) {
            count = 0;
            head = 0;
            gamma = 1.0;
            for (int i = 0; i < n; ++i) d[i] = freev[i] ? -g[i] : 0.0;
        }

        double t = 1.0;
        if (count == 0) {
            const double inv = 1.0 / sqrt(dot(d, d, n));
            t = (inv < 1.0) ? inv : 1.0;
        }

        int accepted = 0;
        double ft = 0.0;
        for (int ls = 0;; ++ls) {
            for (int i = 0; i < n; ++i) xt[i] = clampd(x[i] + t * d[i], lb[i], ub[i]);
            ft = fn(xt, gt, data);
            ++nev;
            double dphi = 0.0;
            for (int i = 0; i < n; ++i) dphi += g[i] * (xt[i] - x[i]);
            if (isfinite(ft) && ft <= f + 1e-4 * dphi) {
                accepted = 1;
                break;
            }
            if (nev >= st->max_evals || ls + 1 >= st->max_ls) break;
            double tn;
            if (dphi < 0.0) {
                const double denom = 2.0 * (ft - f - dphi);
                tn = (-dphi * t) / denom;
            } else {
                tn = 0.5 * t;
            }
            const double lo = 0.1 * t, hi = 0.5 * t;
            tn = (tn >= lo) ? tn : lo; /* also replaces NaN */
            tn = (tn > hi) ? hi : tn;
            t = tn;
        }
        if (!accepted) break;

        /* update: a full ring below M pairs grows (heap) instead of wrapping */
        if (count == hs.cap && hs.cap < M) {
            int want = 4 * hs.cap;
            if (want > M) want = M;
            if (!history_grow(&hs, head, count, want)) abort();
            head = 0;
        }
        double sy = 0.0, yy = 0.0;
        int xstop = 1;
        const int slot = (count < hs.cap) ? (head + count) % hs.cap : head;
        double* Ss = hs.S + (size_t)slot * n;
        double* Ys = hs.Y + (size_t)slot * n;
        for (int i = 0; i < n; ++i) {
            const double s = xt[i] - x[i];
            const double y = gt[i] - g[i];
            sy += s * y;
            yy += y * y;
            xstop = xstop && (fabs(s) < st->xtol_rel * (fabs(xt[i]) + fabs(x[i])) * 0.5 || xt[i] == x[i]);
            Ss[i] = s;
            Ys[i] = y;
            x[i] = xt[i];
            g[i] = gt[i];
        }
        const double fold = f;
        f = ft;
        if (sy > 2.220446049250313e-16 * yy) {
            hs.rh[slot] = 1.0 / sy;
            gamma = sy / yy;
            if (count < hs.cap) ++count;
            else head = (head + 1) % hs.cap;
            if (count > max_run) max_run = count;
        } else {
            count = 0;
            head = 0;
            gamma = 1.0;
        }
        if (nev >= st->max_evals) break;
        if (relstop(fold, f, st->ftol_rel)) break;
        if (xstop) break;
    }
    free(hs.heap);
    if (f_out) *f_out = f;
    if (evals_out) *evals_out = nev;
    if (iterations_out) *iterations_out = iters;
    if (pairs_used_out) *pairs_used_out = used;
    return max_run;
}

typedef struct {
    int L;
    const orc_layer* layers;
    const double* target;
    const double* target_alphas;
    double lambda[ORC_MAX_CONSTRAINTS];
    double rho;
    orc_variant v;
} pixel_problem;

static double pixel_fn(const double* x, double* grad, void* data)
{
    const pixel_problem* p = (const pixel_problem*)data;
    double C[ORC_MAX_CONSTRAINTS];
    return orc_objective_ex(x, p->L, p->layers, p->target, p->target_alphas, p->lambda, p->rho, &p->v, grad, C);
}

static double norm(const double* v, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += v[i] * v[i];
    return sqrt(s);
}

static int solve_pixel_from(const double target[3], int L, const orc_layer* layers, int opaque,
                            const double* init_colors, const double* target_alphas,
                            const double* bg_color, const double* x_init, const orc_settings* st,
                            double* x_out, orc_counters* counters);
static _Thread_local const double* g_lambda_init = NULL;

/* unblending.cpp:155-309 */
int orc_solve_pixel(const double target[3], int L, const orc_layer* layers, int opaque,
                    const double* init_colors, const double* target_alphas,
                    const double* bg_color, const orc_settings* st, double* x_out,
                    orc_counters* counters)
{
    return solve_pixel_from(target, L, layers, opaque, init_colors, target_alphas, bg_color, NULL, st,
                            x_out, counters);
}

/* First pass started from x_init (4L values) instead of the default
 * initialisation; the opaque background still sets alpha_0 = 1. */
int orc_solve_pixel_warm(const double target[3], int L, const orc_layer* layers, int opaque,
                         const double* x_init, const double* lambda_init, const orc_settings* st,
                         double* x_out, orc_counters* counters)
{
    g_lambda_init = lambda_init;
    const int r = solve_pixel_from(target, L, layers, opaque, NULL, NULL, NULL, x_init, st, x_out, counters);
    g_lambda_init = NULL;
    return r;
}

static int solve_pixel_from(const double target[3], int L, const orc_layer* layers, int opaque,
                            const double* init_colors, const double* target_alphas,
                            const double* bg_color, const double* x_init, const orc_settings* st,
                            double* x_out, orc_counters* counters)
{
    const int N = 4 * L;
    const int refine = target_alphas != NULL;
    double lb[ORC_MAX_VARS], ub[ORC_MAX_VARS], x[ORC_MAX_VARS], xn[ORC_MAX_VARS];
    for (int i = 0; i < N; ++i) {
        lb[i] = 0.0;
        ub[i] = 1.0;
    }
    if (refine && st->alpha_fixed)
        for (int i = 0; i < L; ++i) lb[i] = ub[i] = target_alphas[i];

    /* find_initial_solution, :145-149; refinement initial values, :187-191 */
    for (int i = 0; i < L; ++i) {
        x[i] = refine ? target_alphas[i] : 0.50;
        for (int c = 0; c < 3; ++c)
            x[L + 3 * i + c] = refine ? init_colors[3 * i + c] : layers[i].mean[c];
    }
    if (x_init)
        for (int i = 0; i < N; ++i) x[i] = x_init[i];
    if (opaque) { /* :194-198 */
        lb[0] = 1.0;
        if (refine && st->alpha_fixed) ub[0] = 1.0;
        x[0] = 1.0;
    }
    if (bg_color) { /* :201-208 */
        for (int c = 0; c < 3; ++c) lb[L + c] = ub[L + c] = x[L + c] = bg_color[c];
    }

    pixel_problem p;
    p.L = L;
    p.layers = layers;
    p.target = target;
    p.target_alphas = target_alphas;
    /* unblending.cpp:236-243: sparsity in the first pass only */
    p.v.gray_mask = st->gray_mask;
    p.v.sparsity = refine ? 0 : st->sparsity;
    p.v.sigma = st->sigma;
    const int nc = (refine ? 3 + L : 4) + 3 * orc_gray_count(&p.v, L);
    for (int j = 0; j < nc; ++j) p.lambda[j] = g_lambda_init ? g_lambda_init[j] : 0.0;
    p.rho = st->initial_rho;

    int count = 0, capped = 0;
    for (;;) {
        memcpy(xn, x, sizeof(double) * (size_t)N);
        int32_t ev = 0, it = 0;
        int64_t used = 0;
        const int pairs = orc_minimize(xn, lb, ub, N, pixel_fn, &p, st, NULL, &ev, &it, &used);
        if (counters) {
            counters->pairs_used += used;
            counters->evals += ev;
            counters->iterations += it;
            counters->outer += 1;
            if (pairs > counters->max_pairs) counters->max_pairs = pairs;
        }
        double g[ORC_MAX_CONSTRAINTS], gn[ORC_MAX_CONSTRAINTS], dx[ORC_MAX_VARS];
        orc_constraints_ex(x, L, layers, target, target_alphas, &p.v, g);
        orc_constraints_ex(xn, L, layers, target, target_alphas, &p.v, gn);
        for (int j = 0; j < nc; ++j) p.lambda[j] = p.lambda[j] - p.rho * gn[j];
        const double ngn = norm(gn, nc);
        if (ngn > st->gamma * norm(g, nc)) p.rho = p.rho * st->beta;
        for (int i = 0; i < N; ++i) dx[i] = xn[i] - x[i];
        const int unchanged = norm(dx, N) < st->epsilon;
        const int satisfied = ngn < st->epsilon;
        memcpy(x, xn, sizeof(double) * (size_t)N);
        if (unchanged && satisfied) break;
        if (count > st->max_count) {
            capped = 1;
            break;
        }
        ++count;
    }
    memcpy(x_out, x, sizeof(double) * (size_t)N);
    if (counters) counters->pixels += 1;
    return capped;
}
