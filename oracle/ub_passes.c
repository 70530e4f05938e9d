/*
 * ub_passes.c -- the two passes of the oracle over whole images (test
 * infrastructure and CPU baseline).
 *
 * unmix: unblending/src/unblending.cpp:441-475 (per-pixel solve, default init).
 * refine targets: unblending.cpp:351-399 (guided filter of every alpha, crop,
 *   normalise :311-333, optional smoothed background).
 * refine: unblending.cpp:401-436 (per-pixel re-solve).
 * Output packing: unblending.cpp:561-573 ([layer][row][col][RGBA]).
 * Threads stand in for parallelutil::parallel_for_2d: rows are handed out
 * dynamically; results do not depend on the thread count.
 */
#include "ub_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef struct {
    const double* rgb;
    size_t stride;
    int w, h, L;
    const orc_layer* layers;
    int opaque;
    const orc_settings* st;
    const double* first;   /* refine: packed first pass */
    const double* targets; /* refine: [L][n] */
    const double* bg;      /* refine smooth: [3][n] or NULL */
    double* out;
    int next_row;
    pthread_mutex_t lock;
    orc_counters totals;
} pass_job;

static void* pass_worker(void* arg)
{
    pass_job* job = (pass_job*)arg;
    const size_t n = (size_t)job->w * job->h;
    const int L = job->L;
    orc_counters local = {0, 0, 0, 0, 0, 0};
    for (;;) {
        pthread_mutex_lock(&job->lock);
        const int y = job->next_row++;
        pthread_mutex_unlock(&job->lock);
        if (y >= job->h) break;
        for (int x = 0; x < job->w; ++x) {
            const size_t i = (size_t)y * job->w + x;
            const double* target = job->rgb + (size_t)y * job->stride + 3 * (size_t)x;
            double sol[ORC_MAX_VARS];
            if (!job->first) {
                orc_solve_pixel(target, L, job->layers, job->opaque, NULL, NULL, NULL, job->st, sol, &local);
            } else {
                double colors[3 * ORC_MAX_LAYERS], talpha[ORC_MAX_LAYERS], bgc[3];
                for (int l = 0; l < L; ++l) {
                    const double* px = job->first + 4 * ((size_t)l * n + i);
                    for (int c = 0; c < 3; ++c) colors[3 * l + c] = px[c];
                    talpha[l] = job->targets[(size_t)l * n + i];
                }
                if (job->bg) {
                    for (int c = 0; c < 3; ++c) bgc[c] = job->bg[(size_t)c * n + i];
                    for (int c = 0; c < 3; ++c) colors[c] = bgc[c];
                }
                orc_solve_pixel(target, L, job->layers, job->opaque, colors, talpha,
                                job->bg ? bgc : NULL, job->st, sol, &local);
            }
            for (int l = 0; l < L; ++l) {
                double* px = job->out + 4 * ((size_t)l * n + i);
                px[0] = sol[L + 3 * l];
                px[1] = sol[L + 3 * l + 1];
                px[2] = sol[L + 3 * l + 2];
                px[3] = sol[l];
            }
        }
    }
    pthread_mutex_lock(&job->lock);
    job->totals.evals += local.evals;
    job->totals.iterations += local.iterations;
    job->totals.outer += local.outer;
    job->totals.pixels += local.pixels;
    job->totals.pairs_used += local.pairs_used;
    if (local.max_pairs > job->totals.max_pairs) job->totals.max_pairs = local.max_pairs;
    pthread_mutex_unlock(&job->lock);
    return NULL;
}

static void run_pass(pass_job* job, int threads, orc_counters* counters)
{
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    job->next_row = 0;
    memset(&job->totals, 0, sizeof job->totals);
    pthread_mutex_init(&job->lock, NULL);
    pthread_t tid[256];
    for (int t = 1; t < threads; ++t) pthread_create(&tid[t], NULL, pass_worker, job);
    pass_worker(job);
    for (int t = 1; t < threads; ++t) pthread_join(tid[t], NULL);
    pthread_mutex_destroy(&job->lock);
    if (counters) {
        counters->evals += job->totals.evals;
        counters->iterations += job->totals.iterations;
        counters->outer += job->totals.outer;
        counters->pixels += job->totals.pixels;
        counters->pairs_used += job->totals.pairs_used;
        if (job->totals.max_pairs > counters->max_pairs) counters->max_pairs = job->totals.max_pairs;
    }
}

int orc_unmix(const double* rgb, size_t stride, int w, int h, const orc_layer* layers, int L,
              int opaque, const orc_settings* settings, int threads, double* out,
              orc_counters* counters)
{
    if (L < 1 || L > ORC_MAX_LAYERS) return 1;
    orc_settings def;
    if (!settings) {
        orc_default_settings(&def);
        settings = &def;
    }
    pass_job job;
    memset(&job, 0, sizeof job);
    job.rgb = rgb;
    job.stride = stride ? stride : 3 * (size_t)w;
    job.w = w;
    job.h = h;
    job.L = L;
    job.layers = layers;
    job.opaque = opaque;
    job.st = settings;
    job.out = out;
    run_pass(&job, threads, counters);
    return 0;
}

static double crop(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

static void filter_plane(const double* plane, const double* rgb, size_t stride, int w, int h, int radius,
                         double eps, int filter, const double* bounds, int index, double* out)
{
    if (filter == ORC_FILTER_EXACT)
        orc_guided_filter_exact(plane, rgb, stride, w, h, radius, eps, bounds ? bounds[0] : 0.0,
                                bounds ? bounds[1 + index] : 0.0, out);
    else
        orc_guided_filter(plane, rgb, stride, w, h, radius, eps, filter == ORC_FILTER_BRUTE, out);
}

int orc_refine_targets(const double* rgb, size_t stride, int w, int h, const orc_layer* layers,
                       int L, const double* first, int opaque, int smooth, int radius,
                       double eps, int filter, const double* bounds, double* targets, double* bg)
{
    (void)opaque;
    const size_t n = (size_t)w * h;
    if (radius < 0) radius = orc_default_radius(w, h);
    if (eps <= 0.0) eps = 1e-4;
    int all_plus = 1, all_so = 1;
    for (int l = 0; l < L; ++l) {
        const orc_layer* ly = &layers[l];
        all_plus = all_plus && ly->comp_x == 2 && ly->comp_y == 1 && ly->comp_z == 1;
        all_so = all_so && ly->comp_x == 1 && ly->comp_y == 1 && ly->comp_z == 1;
    }
    double* plane = (double*)malloc(sizeof(double) * n);
    /* unblending.cpp:357-363: filter every alpha, layer 0 included */
    for (int l = 0; l < L; ++l) {
        for (size_t i = 0; i < n; ++i) plane[i] = first[4 * ((size_t)l * n + i) + 3];
        filter_plane(plane, rgb, stride, w, h, radius, eps, filter, bounds, l, targets + (size_t)l * n);
    }
    /* :366-389 crop then normalise */
    for (size_t i = 0; i < n; ++i) {
        double a[ORC_MAX_LAYERS];
        for (int l = 0; l < L; ++l) a[l] = crop(targets[(size_t)l * n + i]);
        if (all_so && fabs(a[0] - 1.0) < 1e-05) {
            /* keep */
        } else if (all_plus) {
            double sum = 0.0;
            for (int l = 0; l < L; ++l) sum += a[l];
            for (int l = 0; l < L; ++l) a[l] = a[l] / sum;
        } else {
            for (int l = 0; l < L; ++l) a[l] = 0.0;
        }
        for (int l = 0; l < L; ++l) targets[(size_t)l * n + i] = a[l];
    }
    /* :392-399 smoothed background, used cropped (:413-416, :428) */
    if (smooth && bg) {
        for (int c = 0; c < 3; ++c) {
            for (size_t i = 0; i < n; ++i) plane[i] = first[4 * i + c];
            filter_plane(plane, rgb, stride, w, h, radius, eps, filter, bounds, L + c, bg + (size_t)c * n);
            for (size_t i = 0; i < n; ++i) bg[(size_t)c * n + i] = crop(bg[(size_t)c * n + i]);
        }
    }
    free(plane);
    return 0;
}

int orc_refine(const double* rgb, size_t stride, int w, int h, const orc_layer* layers, int L,
               const double* first, int opaque, int smooth, int radius, double eps, int filter,
               const double* bounds, const orc_settings* settings, int threads, double* out,
               orc_counters* counters)
{
    if (L < 1 || L > ORC_MAX_LAYERS) return 1;
    if (smooth && !opaque) return 1; /* unblending.cpp:599-600 */
    orc_settings def;
    if (!settings) {
        orc_default_settings(&def);
        settings = &def;
    }
    const size_t n = (size_t)w * h;
    double* targets = (double*)malloc(sizeof(double) * n * (size_t)L);
    double* bg = smooth ? (double*)malloc(sizeof(double) * n * 3) : NULL;
    orc_refine_targets(rgb, stride, w, h, layers, L, first, opaque, smooth, radius, eps, filter, bounds,
                       targets, bg);
    /* the first pass may alias out: read it before the solve writes */
    double* first_copy = NULL;
    if (first == out) {
        first_copy = (double*)malloc(sizeof(double) * 4 * n * (size_t)L);
        memcpy(first_copy, first, sizeof(double) * 4 * n * (size_t)L);
        first = first_copy;
    }
    pass_job job;
    memset(&job, 0, sizeof job);
    job.rgb = rgb;
    job.stride = stride ? stride : 3 * (size_t)w;
    job.w = w;
    job.h = h;
    job.L = L;
    job.layers = layers;
    job.opaque = opaque;
    job.st = settings;
    job.first = first;
    job.targets = targets;
    job.bg = bg;
    job.out = out;
    run_pass(&job, threads, counters);
    free(first_copy);
    free(targets);
    free(bg);
    return 0;
}

void orc_recomposite(const double* packed, int w, int h, const orc_layer* layers, int L,
                     double* rgba)
{
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i) {
        double x[ORC_MAX_VARS];
        for (int l = 0; l < L; ++l) {
            const double* px = packed + 4 * ((size_t)l * n + i);
            x[l] = px[3];
            for (int c = 0; c < 3; ++c) x[L + 3 * l + c] = px[c];
        }
        orc_composite_stack(x, L, layers, rgba + 4 * i);
    }
}
