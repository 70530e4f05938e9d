// ub_device.hpp -- per-pixel mathematics for the gfx950 kernels.
//
// The operations follow, expression by expression:
//   blend functions            unblending/include/unblending/blend_mode.hpp:88-202
//   two-layer composite        unblending/src/equations.cpp:10-32 (stack fold :34-74)
//   Gaussian distance/gradient unblending/include/unblending/color_model.hpp:29-37
//   energy, constraints        unblending/src/equations.cpp:88-155
//   gradient dE + J(rho C - l) unblending/src/equations.cpp:157-401, unblending.cpp:67-114
// with the constraint part of the gradient accumulated by a reverse sweep
// over the per-step 4x4 blocks of equations.cpp:222-282.  Everything is FP64
// and is compiled with -ffp-contract=off: each sum is written in the order
// the host implementation of the same formulas uses, so results agree with it
// bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace ubhip {

// A layer as the kernels read it: comp-op factors as doubles, mean, row-major
// precision, blend mode.  16 doubles = 128 bytes.
struct DevLayer {
    double X, Y, Z;
    double mean[3];
    double P[9];
    int mode;
    int pad;
};
static_assert(sizeof(DevLayer) == 128, "DevLayer is 128 bytes");

enum : int {
    kNormal = 0, kMultiply, kScreen, kOverlay, kDarken, kLighten, kColorDodge,
    kColorBurn, kHardLight, kSoftLight, kDifference, kExclusion, kLinearDodge
};

constexpr double kBlendEps = 1e-05;  // blend_mode.hpp:88

__device__ __forceinline__ double blend(double s, double d, int mode)
{
    switch (mode) {
    case kNormal: return s;
    case kMultiply: return s * d;
    case kScreen: return 1.0 - (1.0 - s) * (1.0 - d);
    case kOverlay: return (d <= 0.5) ? 2.0 * s * d : 1.0 - 2.0 * (1.0 - s) * (1.0 - d);
    case kDarken: return (s < d) ? s : d;
    case kLighten: return (s < d) ? d : s;
    case kColorDodge: {
        if (d < kBlendEps) return 0.0;
        if (1.0 - s < kBlendEps) return 1.0;
        const double q = d / (1.0 - s);
        return (q < 1.0) ? q : 1.0;
    }
    case kColorBurn: {
        if (1.0 - d < kBlendEps) return 1.0;
        if (s < kBlendEps) return 0.0;
        const double q = (1.0 - d) / s;
        return 1.0 - ((q < 1.0) ? q : 1.0);
    }
    case kHardLight: return (s <= 0.5) ? 2.0 * s * d : 1.0 - 2.0 * (1.0 - s) * (1.0 - d);
    case kSoftLight: {
        if (s <= 0.5) return d - (1.0 - 2.0 * s) * d * (1.0 - d);
        const double g = (d <= 0.25) ? ((16.0 * d - 12.0) * d + 4.0) * d : __builtin_sqrt(d);
        return d + (2.0 * s - 1.0) * (g - d);
    }
    case kDifference: return (s < d) ? d - s : s - d;
    case kExclusion: return s + d - 2.0 * s * d;
    case kLinearDodge: return s + d;
    default: return 0.0;
    }
}

// Blend value and both partial derivatives in one pass (the three switch
// statements of blend_mode.hpp:90-196 share their branch conditions).
__device__ __forceinline__ void blend_all(double s, double d, int mode, double& D, double& gs, double& gd)
{
    switch (mode) {
    case kNormal: D = s; gs = 1.0; gd = 0.0; return;
    case kMultiply: D = s * d; gs = d; gd = s; return;
    case kScreen: D = 1.0 - (1.0 - s) * (1.0 - d); gs = 1.0 - d; gd = 1.0 - s; return;
    case kOverlay:
        if (d <= 0.5) { D = 2.0 * s * d; gs = 2.0 * d; gd = 2.0 * s; }
        else { D = 1.0 - 2.0 * (1.0 - s) * (1.0 - d); gs = 2.0 * (1.0 - d); gd = 2.0 * (1.0 - s); }
        return;
    case kDarken:
        if (s < d) { D = s; gs = 1.0; gd = 0.0; } else { D = d; gs = 0.0; gd = 1.0; }
        return;
    case kLighten:
        if (s < d) { D = d; gs = 0.0; gd = 1.0; } else { D = s; gs = 1.0; gd = 0.0; }
        return;
    case kColorDodge:
        if (d < kBlendEps) { D = 0.0; gs = 0.0; gd = 0.0; return; }
        if (1.0 - s < kBlendEps) { D = 1.0; gs = 0.0; gd = 0.0; return; }
        {
            const double q = d / (1.0 - s);
            D = (q < 1.0) ? q : 1.0;
            if (1.0 < q) { gs = 0.0; gd = 0.0; }
            else { gs = d / ((1.0 - s) * (1.0 - s)); gd = 1.0 / (1.0 - s); }
        }
        return;
    case kColorBurn:
        if (1.0 - d < kBlendEps) { D = 1.0; gs = 0.0; gd = 0.0; return; }
        if (s < kBlendEps) { D = 0.0; gs = 0.0; gd = 0.0; return; }
        {
            const double q = (1.0 - d) / s;
            D = 1.0 - ((q < 1.0) ? q : 1.0);
            if (1.0 < q) { gs = -0.0; gd = -0.0; }
            else { gs = -(-(1.0 - d) / (s * s)); gd = -(-1.0 / s); }
        }
        return;
    case kHardLight:
        if (s <= 0.5) { D = 2.0 * s * d; gs = 2.0 * d; gd = 2.0 * s; }
        else { D = 1.0 - 2.0 * (1.0 - s) * (1.0 - d); gs = 2.0 * (1.0 - d); gd = 2.0 * (1.0 - s); }
        return;
    case kSoftLight:
        if (s <= 0.5) {
            D = d - (1.0 - 2.0 * s) * d * (1.0 - d);
            gs = 2.0 * d * (1.0 - d);
            gd = 2.0 * s + 2.0 * d - 4.0 * s * d;
        } else {
            double g, h;
            if (d <= 0.25) {
                g = ((16.0 * d - 12.0) * d + 4.0) * d;
                h = 48.0 * d * d - 24.0 * d + 4.0;
            } else {
                const double r = __builtin_sqrt(d);
                g = r;
                h = 1.0 / (2.0 * r);
            }
            D = d + (2.0 * s - 1.0) * (g - d);
            gs = 2.0 * (g - d);
            gd = 1.0 + (2.0 * s - 1.0) * h - (2.0 * s - 1.0);
        }
        return;
    case kDifference:
        if (s < d) { D = d - s; gs = -1.0; gd = 1.0; } else { D = s - d; gs = 1.0; gd = -1.0; }
        return;
    case kExclusion: D = s + d - 2.0 * s * d; gs = 1.0 - 2.0 * d; gd = 1.0 - 2.0 * s; return;
    case kLinearDodge: D = s + d; gs = 1.0; gd = 1.0; return;
    default: D = 0.0; gs = 0.0; gd = 0.0; return;
    }
}

// (c - mu)^T P (c - mu), row vector times matrix first (color_model.hpp:29-31).
__device__ __forceinline__ double distance(const double c0, const double c1, const double c2,
                                           const DevLayer& l)
{
    const double d0 = c0 - l.mean[0], d1 = c1 - l.mean[1], d2 = c2 - l.mean[2];
    const double t0 = d0 * l.P[0] + d1 * l.P[3] + d2 * l.P[6];
    const double t1 = d0 * l.P[1] + d1 * l.P[4] + d2 * l.P[7];
    const double t2 = d0 * l.P[2] + d1 * l.P[5] + d2 * l.P[8];
    return t0 * d0 + t1 * d1 + t2 * d2;
}

// 2 P (c - mu) (color_model.hpp:34-37).
__device__ __forceinline__ void distance_gradient(const double c0, const double c1, const double c2,
                                                  const DevLayer& l, double g[3])
{
    const double d0 = c0 - l.mean[0], d1 = c1 - l.mean[1], d2 = c2 - l.mean[2];
#pragma unroll
    for (int j = 0; j < 3; ++j)
        g[j] = (2.0 * l.P[3 * j + 0]) * d0 + (2.0 * l.P[3 * j + 1]) * d1 + (2.0 * l.P[3 * j + 2]) * d2;
}

__device__ __forceinline__ int uniform_mode(int mode)
{
    return __builtin_amdgcn_readfirstlane(mode);
}

// Number of constraints: 4 in the first pass, 3 + L in refinement.
template <int L, bool REFINE>
struct Shape {
    static constexpr int N = 4 * L;
    static constexpr int NC = REFINE ? 3 + L : 4;
};

// Augmented-Lagrangian objective, its gradient and the constraint vector at x.
// x = [alpha_0..alpha_{L-1}, c_0 rgb .. c_{L-1} rgb].
template <int L, bool REFINE>
__device__ __forceinline__ double objective(const double* __restrict__ x, const DevLayer* lay,
                                            const double tgt[3], const double* ta,
                                            const double* lam, const double rho,
                                            double* __restrict__ grad, double* __restrict__ C)
{
    constexpr int NC = Shape<L, REFINE>::NC;
    const double* alphas = x;
    const double* colors = x + L;

    double pre[L][4];
    pre[0][0] = colors[0];
    pre[0][1] = colors[1];
    pre[0][2] = colors[2];
    pre[0][3] = alphas[0];
#pragma unroll
    for (int k = 1; k < L; ++k) {
        const DevLayer& l = lay[k];
        const int mode = uniform_mode(l.mode);
        const double as = alphas[k], ad = pre[k - 1][3];
        const double a = l.X * as * ad + l.Y * as * (1.0 - ad) + l.Z * ad * (1.0 - as);
        const double ky = l.Y * as * (1.0 - ad);
        const double kz = l.Z * ad * (1.0 - as);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double cs = colors[3 * k + c], cd = pre[k - 1][c];
            const double f = blend(cs, cd, mode);
            const double p = f * as * ad + ky * cs + kz * cd;
            pre[k][c] = (a > 1e-12) ? p / a : p;
        }
        pre[k][3] = a;
    }

#pragma unroll
    for (int c = 0; c < 3; ++c) C[c] = pre[L - 1][c] - tgt[c];
    if constexpr (REFINE) {
#pragma unroll
        for (int i = 0; i < L; ++i) C[3 + i] = alphas[i] - ta[i];
    } else {
        C[3] = pre[L - 1][3] - 1.0;
    }

    double energy = 0.0;
    double dist[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        dist[i] = distance(colors[3 * i], colors[3 * i + 1], colors[3 * i + 2], lay[i]);
        energy += alphas[i] * dist[i];
    }
    if constexpr (!REFINE) {
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < L; ++i) sum += alphas[i];
        energy += 0.01 * sum;
    }
    double lagrange = 0.0, squared = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) lagrange += lam[j] * C[j];
#pragma unroll
    for (int j = 0; j < NC; ++j) squared += C[j] * C[j];
    const double f = energy + (-lagrange) + 0.5 * rho * squared;

    double w[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) w[j] = rho * C[j] - lam[j];

    double aB0 = w[0], aB1 = w[1], aB2 = w[2];
    double aA = REFINE ? 0.0 : w[3 < NC ? 3 : 0];
#pragma unroll
    for (int k = L - 1; k >= 1; --k) {
        const DevLayer& l = lay[k];
        const int mode = uniform_mode(l.mode);
        const double as = alphas[k], ad = pre[k - 1][3];
        const double A = pre[k][3];
        const double pA_s = l.X * ad + l.Y * (1.0 - ad) - l.Z * ad;
        const double pA_d = l.X * as - l.Y * as + l.Z * (1.0 - as);
        double adj_as = 0.0, adj_ad = 0.0;
        double aB[3] = {aB0, aB1, aB2};
        double nB[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double cs = colors[3 * k + c], cd = pre[k - 1][c];
            double D, gs, gd;
            blend_all(cs, cd, mode, D, gs, gd);
            const double B = pre[k][c];
            const double dB_dcs = (as * ad * gs + l.Y * (1.0 - ad) * as) / A;
            const double dB_das = ((D * ad + l.Y * (1.0 - ad) * cs - l.Z * ad * cd) - B * pA_s) / A;
            const double dB_dcd = (as * ad * gd + l.Z * (1.0 - as) * ad) / A;
            const double dB_dad = ((D * as - l.Y * as * cs + l.Z * (1.0 - as) * cd) - B * pA_d) / A;
            grad[L + 3 * k + c] = dB_dcs * aB[c];
            adj_as += dB_das * aB[c];
            adj_ad += dB_dad * aB[c];
            nB[c] = dB_dcd * aB[c];
        }
        grad[k] = adj_as + pA_s * aA;
        aA = adj_ad + pA_d * aA;
        aB0 = nB[0];
        aB1 = nB[1];
        aB2 = nB[2];
    }
    grad[0] = aA;
    grad[L + 0] = aB0;
    grad[L + 1] = aB1;
    grad[L + 2] = aB2;
    if constexpr (REFINE) {
#pragma unroll
        for (int i = 0; i < L; ++i) grad[i] += w[3 + i];
    }
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const double ge = REFINE ? dist[i] : 0.01;
        grad[i] = ge + grad[i];
        double g[3];
        distance_gradient(colors[3 * i], colors[3 * i + 1], colors[3 * i + 2], lay[i], g);
#pragma unroll
        for (int j = 0; j < 3; ++j) grad[L + 3 * i + j] = alphas[i] * g[j] + grad[L + 3 * i + j];
    }
    return f;
}

}  // namespace ubhip
