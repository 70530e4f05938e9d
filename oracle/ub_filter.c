/*
 * ub_filter.c -- box and colour guided filter of the oracle (test infrastructure).
 *
 * Restates unblending/src/image_processing.cpp:120-151 (clamp-edge convolution),
 * unblending/include/unblending/image_processing.hpp:215-222 (box filter) and
 * unblending/src/image_processing.cpp:214-273 (He et al. colour guided filter,
 * with its statistics, inverse and summation order).
 */
#include "ub_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* image_processing.hpp:215-222 and image_processing.cpp:120-151 */
static void box_brute(const double* in, int w, int h, int r, double* out)
{
    const int size = 2 * r + 1;
    const double weight = 1.0 / (double)(size * size);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double value = 0.0;
            for (int kx = 0; kx < size; ++kx)
                for (int ky = 0; ky < size; ++ky) {
                    const int ox = clampi(x + kx - r, 0, w - 1);
                    const int oy = clampi(y + ky - r, 0, h - 1);
                    value += weight * in[(size_t)oy * w + ox];
                }
            out[(size_t)y * w + x] = value;
        }
}

/* The same clamp-edge box as separable running window sums. */
static void box_sliding(const double* in, int w, int h, int r, double* out)
{
    const int size = 2 * r + 1;
    const double weight = 1.0 / (double)(size * size);
    double* tmp = (double*)malloc(sizeof(double) * (size_t)w * h);
    for (int y = 0; y < h; ++y) {
        const double* row = in + (size_t)y * w;
        double s = 0.0;
        for (int dx = -r; dx <= r; ++dx) s += row[clampi(dx, 0, w - 1)];
        tmp[(size_t)y * w] = s;
        for (int x = 1; x < w; ++x) {
            s += row[clampi(x + r, 0, w - 1)] - row[clampi(x - r - 1, 0, w - 1)];
            tmp[(size_t)y * w + x] = s;
        }
    }
    for (int x = 0; x < w; ++x) {
        double s = 0.0;
        for (int dy = -r; dy <= r; ++dy) s += tmp[(size_t)clampi(dy, 0, h - 1) * w + x];
        out[x] = s * weight;
        for (int y = 1; y < h; ++y) {
            s += tmp[(size_t)clampi(y + r, 0, h - 1) * w + x] - tmp[(size_t)clampi(y - r - 1, 0, h - 1) * w + x];
            out[(size_t)y * w + x] = s * weight;
        }
    }
    free(tmp);
}

void orc_box(const double* in, int w, int h, int r, int brute, double* out)
{
    if (r <= 0) { /* image_processing.hpp:218 */
        memcpy(out, in, sizeof(double) * (size_t)w * h);
        return;
    }
    if (brute) box_brute(in, w, h, r, out);
    else box_sliding(in, w, h, r, out);
}

int orc_default_radius(int w, int h)
{
    return 60 * (w < h ? w : h) / 1000; /* unblending.cpp:351 */
}

/* image_processing.cpp:214-273 */
void orc_guided_filter(const double* p, const double* rgb, size_t stride, int w, int h, int r,
                       double eps, int brute, double* q)
{
    const size_t n = (size_t)w * h;
    if (!stride) stride = 3 * (size_t)w;
    double* buf = (double*)malloc(sizeof(double) * n * 32);
    double* I[3] = {buf, buf + n, buf + 2 * n};
    double* mI[3] = {buf + 3 * n, buf + 4 * n, buf + 5 * n};
    double* mp = buf + 6 * n;
    double* cov[3] = {buf + 7 * n, buf + 8 * n, buf + 9 * n};
    double* var[6] = {buf + 10 * n, buf + 11 * n, buf + 12 * n, buf + 13 * n, buf + 14 * n, buf + 15 * n};
    double* a[3] = {buf + 16 * n, buf + 17 * n, buf + 18 * n};
    double* b = buf + 19 * n;
    double* t0 = buf + 20 * n;
    double* t1 = buf + 21 * n;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) I[c][(size_t)y * w + x] = rgb[(size_t)y * stride + 3 * (size_t)x + c];

    for (int c = 0; c < 3; ++c) orc_box(I[c], w, h, r, brute, mI[c]);
    orc_box(p, w, h, r, brute, mp);
    for (int c = 0; c < 3; ++c) {
        for (size_t i = 0; i < n; ++i) t0[i] = I[c][i] * p[i];
        orc_box(t0, w, h, r, brute, t1);
        for (size_t i = 0; i < n; ++i) cov[c][i] = t1[i] - mI[c][i] * mp[i];
    }
    static const int pairs[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int k = 0; k < 6; ++k) {
        const int u = pairs[k][0], v = pairs[k][1];
        for (size_t i = 0; i < n; ++i) t0[i] = I[u][i] * I[v][i];
        orc_box(t0, w, h, r, brute, t1);
        for (size_t i = 0; i < n; ++i) var[k][i] = t1[i] - mI[u][i] * mI[v][i];
    }
    for (size_t i = 0; i < n; ++i) {
        /* (Sigma + eps I)^-1 by cofactors (Eigen's closed-form 3x3 inverse) */
        const double m00 = var[0][i] + eps, m01 = var[1][i], m02 = var[2][i];
        const double m11 = var[3][i] + eps, m12 = var[4][i], m22 = var[5][i] + eps;
        const double c00 = m11 * m22 - m12 * m12;
        const double c01 = m02 * m12 - m01 * m22;
        const double c02 = m01 * m12 - m02 * m11;
        const double c11 = m00 * m22 - m02 * m02;
        const double c12 = m01 * m02 - m00 * m12;
        const double c22 = m00 * m11 - m01 * m01;
        const double det = m00 * c00 + m01 * c01 + m02 * c02;
        const double inv = 1.0 / det;
        const double v0 = cov[0][i], v1 = cov[1][i], v2 = cov[2][i];
        a[0][i] = (c00 * inv) * v0 + (c01 * inv) * v1 + (c02 * inv) * v2;
        a[1][i] = (c01 * inv) * v0 + (c11 * inv) * v1 + (c12 * inv) * v2;
        a[2][i] = (c02 * inv) * v0 + (c12 * inv) * v1 + (c22 * inv) * v2;
        b[i] = ((mp[i] - a[0][i] * mI[0][i]) - a[1][i] * mI[1][i]) - a[2][i] * mI[2][i];
    }
    orc_box(b, w, h, r, brute, q);
    for (int c = 0; c < 3; ++c) {
        orc_box(a[c], w, h, r, brute, t1);
        for (size_t i = 0; i < n; ++i) q[i] = q[i] + t1[i] * I[c][i];
    }
    free(buf);
}

/* ---- exact-sum guided filter (the device filter's definition) ----
 *
 * Every box mean is taken over values quantised to a grid 2^-k:
 *   Q(v) = rint(v 2^k),  mean = ldexp(sum Q, -k) * w,  w = 1 / (2r + 1)^2,
 * with k chosen per quantity from a bound |v| <= 2^E so that every partial sum
 * the device forms stays within 2^53: the device sums the integers Q in FP64,
 * where such sums are exact, in whatever order its tiling gives; here they are
 * summed in 64-bit integers.  Either way the statistics do not depend on the
 * order of summation.  The device's longest partial sums are prefixes over
 * P(r) columns of (2r + 1)-row window sums, P(r) = 256 for r <= 96 and 1024
 * beyond, so k = 53 - E - ceil(log2(P(r) (2r + 1))).  The per-pixel
 * arithmetic after the sums is the reference's (image_processing.cpp:226-270).
 * Bounds: guidance statistics (I, I I) from max|I|; a channel's (p, I p) from
 * max|I| and max|p|; the coefficients from |a_c| <= max|p| / sqrt(eps) and
 * |b| <= max|p| + s3 max|I| max|p| / sqrt(eps), s3 = sqrt(3) rounded up
 * (twice the bounds |a|_2 <= sd(p) / (2 sqrt(eps)), which hold because the
 * statistics are, up to the grid step, covariances of one data set, and
 * sd(p) <= max|p|).  Quantised values are clamped to the bound, so the sums
 * stay exact whatever the data.
 * r = 0 keeps the reference's identity box (image_processing.hpp:218). */

int orc_qexp(double x)
{
    uint64_t b;
    x = fabs(x);
    memcpy(&b, &x, sizeof b);
    const int ex = (int)((b >> 52) & 0x7ff);
    return ex == 0 ? -1022 : ex - 1022; /* smallest e with |x| < 2^e (zero: -1022) */
}

int orc_qprefix(int r)
{
    return r <= 96 ? 256 : 1024;
}

int orc_qscale(int e, int r)
{
    const long long n = (long long)orc_qprefix(r) * (2 * r + 1);
    int c = 0;
    while ((1LL << c) < n) ++c;
    int k = 53 - e - c;
    return k > 1000 ? 1000 : k;
}

/* the coefficient bounds of a channel: |a_c| <= ba, |b| <= bb */
void orc_coeff_bounds(double imax, double pmax, double eps, double* ba, double* bb)
{
    const double isr = 1.0 / sqrt(eps);
    *ba = pmax * isr;
    *bb = pmax + (ORC_SQRT3_UP * imax) * *ba;
}

static int64_t quantise(double v, int k, int e)
{
    double t = rint(ldexp(v, k));
    const double lim = ldexp(1.0, e + k); /* the bound, on the grid; never reached by bounded values */
    t = t > lim ? lim : (t < -lim ? -lim : t);
    return (int64_t)t;
}

/* clamp-edge box sums of n-element integer planes: out[i] = sum over the window */
static void box_sum_i64(const int64_t* in, int w, int h, int r, int64_t* out)
{
    int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * (size_t)w * h);
    for (int y = 0; y < h; ++y) {
        const int64_t* row = in + (size_t)y * w;
        int64_t s = 0;
        for (int dx = -r; dx <= r; ++dx) s += row[clampi(dx, 0, w - 1)];
        tmp[(size_t)y * w] = s;
        for (int x = 1; x < w; ++x) {
            s += row[clampi(x + r, 0, w - 1)] - row[clampi(x - r - 1, 0, w - 1)];
            tmp[(size_t)y * w + x] = s;
        }
    }
    for (int x = 0; x < w; ++x) {
        int64_t s = 0;
        for (int dy = -r; dy <= r; ++dy) s += tmp[(size_t)clampi(dy, 0, h - 1) * w + x];
        out[x] = s;
        for (int y = 1; y < h; ++y) {
            s += tmp[(size_t)clampi(y + r, 0, h - 1) * w + x] - tmp[(size_t)clampi(y - r - 1, 0, h - 1) * w + x];
            out[(size_t)y * w + x] = s;
        }
    }
    free(tmp);
}

/* mean of the quantised values of plane v over every window (r > 0), or v itself (r = 0) */
static void qbox(const double* v, int w, int h, int r, int e, double* out)
{
    const size_t n = (size_t)w * h;
    if (r <= 0) {
        memcpy(out, v, sizeof(double) * n);
        return;
    }
    const int k = orc_qscale(e, r);
    const double weight = 1.0 / (double)((2 * r + 1) * (2 * r + 1));
    int64_t* qi = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t* qs = (int64_t*)malloc(sizeof(int64_t) * n);
    for (size_t i = 0; i < n; ++i) qi[i] = quantise(v[i], k, e);
    box_sum_i64(qi, w, h, r, qs);
    for (size_t i = 0; i < n; ++i) out[i] = ldexp((double)qs[i], -k) * weight;
    free(qi);
    free(qs);
}

double orc_max_abs(const double* v, size_t n, size_t step)
{
    double m = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double a = fabs(v[i * step]);
        m = a > m ? a : m;
    }
    return m;
}

void orc_guided_filter_exact(const double* p, const double* rgb, size_t stride, int w, int h, int r, double eps,
                             double imax, double pmax, double* q)
{
    const size_t n = (size_t)w * h;
    if (!stride) stride = 3 * (size_t)w;
    double* buf = (double*)malloc(sizeof(double) * n * 24);
    double* I[3] = {buf, buf + n, buf + 2 * n};
    double* mI[3] = {buf + 3 * n, buf + 4 * n, buf + 5 * n};
    double* mp = buf + 6 * n;
    double* cov[3] = {buf + 7 * n, buf + 8 * n, buf + 9 * n};
    double* var[6] = {buf + 10 * n, buf + 11 * n, buf + 12 * n, buf + 13 * n, buf + 14 * n, buf + 15 * n};
    double* a[3] = {buf + 16 * n, buf + 17 * n, buf + 18 * n};
    double* b = buf + 19 * n;
    double* t0 = buf + 20 * n;
    double* t1 = buf + 21 * n;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) I[c][(size_t)y * w + x] = rgb[(size_t)y * stride + 3 * (size_t)x + c];
    if (!(imax > 0.0)) imax = fmax(fmax(orc_max_abs(I[0], n, 1), orc_max_abs(I[1], n, 1)), orc_max_abs(I[2], n, 1));
    if (!(pmax > 0.0)) pmax = orc_max_abs(p, n, 1);
    const int eI = orc_qexp(imax), eP = orc_qexp(pmax);
    const int eG = eI > 2 * eI ? eI : 2 * eI; /* I and I I */
    const int eC = eP > eI + eP ? eP : eI + eP; /* p and I p */
    double abound, bbound;
    orc_coeff_bounds(imax, pmax, eps, &abound, &bbound);
    const int eA = orc_qexp(abound), eB = orc_qexp(bbound);

    for (int c = 0; c < 3; ++c) qbox(I[c], w, h, r, eG, mI[c]);
    qbox(p, w, h, r, eC, mp);
    for (int c = 0; c < 3; ++c) {
        for (size_t i = 0; i < n; ++i) t0[i] = I[c][i] * p[i];
        qbox(t0, w, h, r, eC, t1);
        for (size_t i = 0; i < n; ++i) cov[c][i] = t1[i] - mI[c][i] * mp[i];
    }
    static const int pairs[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int k = 0; k < 6; ++k) {
        const int u = pairs[k][0], v = pairs[k][1];
        for (size_t i = 0; i < n; ++i) t0[i] = I[u][i] * I[v][i];
        qbox(t0, w, h, r, eG, t1);
        for (size_t i = 0; i < n; ++i) var[k][i] = t1[i] - mI[u][i] * mI[v][i];
    }
    for (size_t i = 0; i < n; ++i) {
        const double m00 = var[0][i] + eps, m01 = var[1][i], m02 = var[2][i];
        const double m11 = var[3][i] + eps, m12 = var[4][i], m22 = var[5][i] + eps;
        const double c00 = m11 * m22 - m12 * m12;
        const double c01 = m02 * m12 - m01 * m22;
        const double c02 = m01 * m12 - m02 * m11;
        const double c11 = m00 * m22 - m02 * m02;
        const double c12 = m01 * m02 - m00 * m12;
        const double c22 = m00 * m11 - m01 * m01;
        const double det = m00 * c00 + m01 * c01 + m02 * c02;
        const double inv = 1.0 / det;
        const double v0 = cov[0][i], v1 = cov[1][i], v2 = cov[2][i];
        a[0][i] = (c00 * inv) * v0 + (c01 * inv) * v1 + (c02 * inv) * v2;
        a[1][i] = (c01 * inv) * v0 + (c11 * inv) * v1 + (c12 * inv) * v2;
        a[2][i] = (c02 * inv) * v0 + (c12 * inv) * v1 + (c22 * inv) * v2;
        b[i] = ((mp[i] - a[0][i] * mI[0][i]) - a[1][i] * mI[1][i]) - a[2][i] * mI[2][i];
    }
    qbox(b, w, h, r, eB, q);
    for (int c = 0; c < 3; ++c) {
        qbox(a[c], w, h, r, eA, t1);
        for (size_t i = 0; i < n; ++i) q[i] = q[i] + t1[i] * I[c][i];
    }
    free(buf);
}
