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

// Blend value and partial derivatives of one channel, without per-lane
// branches: both sides of every condition of blend_mode.hpp:90-196 are
// evaluated and the result is selected, which yields the same values as the
// branching form (an unselected side may divide by zero; it is discarded).
template <int MODE>
__device__ __forceinline__ void blend_sel(double s, double d, double& D, double& gs, double& gd)
{
    if constexpr (MODE == kNormal) {
        D = s; gs = 1.0; gd = 0.0;
    } else if constexpr (MODE == kMultiply) {
        D = s * d; gs = d; gd = s;
    } else if constexpr (MODE == kScreen) {
        D = 1.0 - (1.0 - s) * (1.0 - d); gs = 1.0 - d; gd = 1.0 - s;
    } else if constexpr (MODE == kOverlay || MODE == kHardLight) {
        const bool lo = (MODE == kOverlay) ? (d <= 0.5) : (s <= 0.5);
        const double D1 = 2.0 * s * d, D2 = 1.0 - 2.0 * (1.0 - s) * (1.0 - d);
        D = lo ? D1 : D2;
        if constexpr (MODE == kOverlay) {
            gs = lo ? 2.0 * d : 2.0 * (1.0 - d);
            gd = lo ? 2.0 * s : 2.0 * (1.0 - s);
        } else {
            gs = lo ? 2.0 * d : 2.0 * (1.0 - d);
            gd = lo ? 2.0 * s : 2.0 * (1.0 - s);
        }
    } else if constexpr (MODE == kDarken) {
        const bool lt = s < d;
        D = lt ? s : d; gs = lt ? 1.0 : 0.0; gd = lt ? 0.0 : 1.0;
    } else if constexpr (MODE == kLighten) {
        const bool lt = s < d;
        D = lt ? d : s; gs = lt ? 0.0 : 1.0; gd = lt ? 1.0 : 0.0;
    } else if constexpr (MODE == kColorDodge) {
        const bool c1 = d < kBlendEps, c2 = 1.0 - s < kBlendEps;
        const double q = d / (1.0 - s);
        const double mid = (q < 1.0) ? q : 1.0;
        const bool flat = 1.0 < q;
        const double gsm = flat ? 0.0 : d / ((1.0 - s) * (1.0 - s));
        const double gdm = flat ? 0.0 : 1.0 / (1.0 - s);
        D = c1 ? 0.0 : (c2 ? 1.0 : mid);
        gs = (c1 || c2) ? 0.0 : gsm;
        gd = (c1 || c2) ? 0.0 : gdm;
    } else if constexpr (MODE == kColorBurn) {
        const bool c1 = 1.0 - d < kBlendEps, c2 = s < kBlendEps;
        const double q = (1.0 - d) / s;
        const double mid = 1.0 - ((q < 1.0) ? q : 1.0);
        const bool flat = 1.0 < q;
        const double gsm = flat ? -0.0 : -(-(1.0 - d) / (s * s));
        const double gdm = flat ? -0.0 : -(-1.0 / s);
        D = c1 ? 1.0 : (c2 ? 0.0 : mid);
        gs = (c1 || c2) ? 0.0 : gsm;
        gd = (c1 || c2) ? 0.0 : gdm;
    } else if constexpr (MODE == kSoftLight) {
        const bool lo = s <= 0.5, small = d <= 0.25;
        const double r = __builtin_sqrt(d);
        const double g = small ? ((16.0 * d - 12.0) * d + 4.0) * d : r;
        const double h = small ? (48.0 * d * d - 24.0 * d + 4.0) : 1.0 / (2.0 * r);
        D = lo ? d - (1.0 - 2.0 * s) * d * (1.0 - d) : d + (2.0 * s - 1.0) * (g - d);
        gs = lo ? 2.0 * d * (1.0 - d) : 2.0 * (g - d);
        gd = lo ? 2.0 * s + 2.0 * d - 4.0 * s * d : 1.0 + (2.0 * s - 1.0) * h - (2.0 * s - 1.0);
    } else if constexpr (MODE == kDifference) {
        const bool lt = s < d;
        D = lt ? d - s : s - d; gs = lt ? -1.0 : 1.0; gd = lt ? 1.0 : -1.0;
    } else if constexpr (MODE == kExclusion) {
        D = s + d - 2.0 * s * d; gs = 1.0 - 2.0 * d; gd = 1.0 - 2.0 * s;
    } else if constexpr (MODE == kLinearDodge) {
        D = s + d; gs = 1.0; gd = 1.0;
    } else {
        D = 0.0; gs = 0.0; gd = 0.0;
    }
}

// The three channels of one layer under one (uniform) mode: a single scalar
// branch per layer, the channels as independent work inside it.  With
// WANT_GRAD false only the blend value is produced.
template <bool WANT_GRAD>
__device__ __forceinline__ void blend3_impl(int mode, const double* s, const double* d, double* D, double* gs,
                                            double* gd)
{
    double tgs[3], tgd[3];
    double* ogs = WANT_GRAD ? gs : tgs;
    double* ogd = WANT_GRAD ? gd : tgd;
#define UB_CASE(M) \
    case M: _Pragma("unroll") for (int c = 0; c < 3; ++c) blend_sel<M>(s[c], d[c], D[c], ogs[c], ogd[c]); return;
    switch (mode) {
        UB_CASE(kNormal)
        UB_CASE(kMultiply)
        UB_CASE(kScreen)
        UB_CASE(kOverlay)
        UB_CASE(kDarken)
        UB_CASE(kLighten)
        UB_CASE(kColorDodge)
        UB_CASE(kColorBurn)
        UB_CASE(kHardLight)
        UB_CASE(kSoftLight)
        UB_CASE(kDifference)
        UB_CASE(kExclusion)
        UB_CASE(kLinearDodge)
    default:
        _Pragma("unroll") for (int c = 0; c < 3; ++c) { D[c] = 0.0; ogs[c] = 0.0; ogd[c] = 0.0; }
        return;
    }
#undef UB_CASE
}

__device__ __forceinline__ void blend3(int mode, const double* s, const double* d, double* D)
{
    blend3_impl<false>(mode, s, d, D, nullptr, nullptr);
}

__device__ __forceinline__ void blend3_all(int mode, const double* s, const double* d, double* D, double* gs,
                                           double* gd)
{
    blend3_impl<true>(mode, s, d, D, gs, gd);
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

// The layer table is uniform and read with scalar loads.  Indexing it through
// an opaque zero keeps the loads from being merged with, or hoisted to, the
// loads of another sweep; otherwise the scheduler loads all layers'
// parameters together and spills the SGPRs.
__device__ __forceinline__ const DevLayer& layer_at(const DevLayer* lay, int k)
{
    int zero;
    asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
    return lay[k + zero];
}

// The same, with the offset also depending on dep: the loads cannot be issued
// before dep exists.  The sweeps pass a value computed early in the step
// before the layer's, so each layer's parameters are loaded one step ahead
// (the scalar-load latency is covered) and not held across earlier steps.
__device__ __forceinline__ const DevLayer& layer_after(const DevLayer* lay, int k, double dep)
{
    int zero;
    asm volatile("s_mov_b32 %0, 0" : "=s"(zero) : "v"(dep));
    return lay[k + zero];
}

// Number of constraints: 4 in the first pass, 3 + L in refinement, and 3
// more for a gray layer (GRAY: layer 1, the reference's SECOND_LAYER_GRAY
// build, unblending.cpp:168-172).
template <int L, bool REFINE, bool GRAY = false>
struct Shape {
    static constexpr int N = 4 * L;
    static constexpr int NB = REFINE ? 3 + L : 4;  // colour and alpha rows
    static constexpr int NC = NB + (GRAY ? 3 : 0);
};

constexpr double kSqrt3 = 1.7320508075688772;  // std::sqrt(3), correctly rounded

// The gray rows of a colour c (equations.cpp:144-152): sqrt(3) c - |c| 1.
__device__ __forceinline__ void gray_rows(const double* c, double* C)
{
    const double norm = __builtin_sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
#pragma unroll
    for (int j = 0; j < 3; ++j) C[j] = kSqrt3 * c[j] - norm * 1.0;
}

// Augmented-Lagrangian objective and its gradient at x.
// x = [alpha_0..alpha_{L-1}, c_0 rgb .. c_{L-1} rgb].  The forward sweep
// composites bottom to top (it reads only the comp-op factors and blend mode
// of each layer); the reverse sweep applies the transposed 4x4 blocks of
// equations.cpp:222-282 (one reciprocal of A per layer) to rho C - lambda,
// evaluates each layer's colour model (distance and its gradient) and writes
// grad (energy and constraint parts together); the energy is then summed
// bottom to top, in the reference's order.  Evaluating the colour models in
// the reverse sweep only keeps each layer's mean and precision out of the
// registers of the forward sweep (same values).  tgt, ta and lam are
// indexable (plain arrays, or per-lane LDS vectors).  Cc receives the
// constraints that depend on the composite: the colour residual (3) and, in
// the first pass, the alpha residual (the refinement's alpha constraints
// alpha_i - t_i need no composite).
//
// Problem variants: GRAY adds layer 1's gray rows to C and their part of
// J (rho C - lambda) -- (sqrt(3) I - c 1^T / |c|) w, zero when |c| <= 1e-3
// (equations.cpp:385-398), each row summed in column order -- to that
// colour's gradient after the sweep; SPARSE adds the sparsity term
// sparse_sigma (sum a / sum a^2 - 1) to the energy (equations.cpp:104-105;
// first pass only, where the alpha block's overwrite discards its gradient).
template <int L, bool REFINE, bool GRAY = false, bool SPARSE = false, class TG, class TA, class TL>
__device__ __forceinline__ double objective(const double* __restrict__ x, const DevLayer* lay, const TG& tgt,
                                            const TA& ta, const TL& lam, const double rho,
                                            double* __restrict__ grad, double* __restrict__ Cc,
                                            const double sparse_sigma = 0.0)
{
    constexpr int NC = Shape<L, REFINE, GRAY>::NC;
    constexpr int NB = Shape<L, REFINE, GRAY>::NB;
    const double* alphas = x;
    const double* colors = x + L;

    double pre[L][4];
    double C[NC];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const DevLayer& l = k < 2 ? layer_at(lay, k) : layer_after(lay, k, pre[k - 2][3]);
        const double* ck = colors + 3 * k;
        if (k == 0) {
            pre[0][0] = ck[0];
            pre[0][1] = ck[1];
            pre[0][2] = ck[2];
            pre[0][3] = alphas[0];
        } else {
            const int mode = uniform_mode(l.mode);
            const double as = alphas[k], ad = pre[k - 1][3];
            const double a = l.X * as * ad + l.Y * as * (1.0 - ad) + l.Z * ad * (1.0 - as);
            const double ky = l.Y * as * (1.0 - ad);
            const double kz = l.Z * ad * (1.0 - as);
            double f[3];
            blend3(mode, ck, pre[k - 1], f);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double p = f[c] * as * ad + ky * ck[c] + kz * pre[k - 1][c];
                pre[k][c] = (a > 1e-12) ? p / a : p;
            }
            pre[k][3] = a;
        }
    }

#pragma unroll
    for (int c = 0; c < 3; ++c) C[c] = pre[L - 1][c] - tgt[c];
    if constexpr (REFINE) {
#pragma unroll
        for (int i = 0; i < L; ++i) C[3 + i] = alphas[i] - ta[i];
    } else {
        C[3] = pre[L - 1][3] - 1.0;
    }
    if constexpr (GRAY) gray_rows(colors + 3, C + NB);
#pragma unroll
    for (int j = 0; j < (REFINE ? 3 : 4); ++j) Cc[j] = C[j];
    double lagrange = 0.0, squared = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) lagrange += lam[j] * C[j];
#pragma unroll
    for (int j = 0; j < NC; ++j) squared += C[j] * C[j];

    // rho C - lambda: the composite rows here; the refinement's alpha rows
    // (rho (alpha_i - t_i) - lambda_{3+i}, the same expression) where the
    // reverse sweep uses them, and the gray rows after it, so that they are
    // not live across it
    constexpr int NW = REFINE ? 3 : 4;
    double w[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) w[j] = rho * C[j] - lam[j];

    double aB[3] = {w[0], w[1], w[2]};
    double aA = REFINE ? 0.0 : w[3 < NW ? 3 : 0];
    double dep = w[0], aprev = w[0];
    double dist[L];  // the layers' colour-model distances (the energy's terms)
#pragma unroll
    for (int k = L - 1; k >= 1; --k) {
        const DevLayer& l = layer_after(lay, k, dep);
        // refinement: the blend mode is read once the reverse chain has left
        // the step two layers up, so the blend (a branch on the mode) is not
        // hoisted, with its operands, ahead of earlier layers' steps (the
        // first-pass kernel is faster without this; measured)
        const int mode = uniform_mode(REFINE ? layer_after(lay, k, aprev).mode : l.mode);
        const double* cs = colors + 3 * k;
        const double* cd = pre[k - 1];
        const double as = alphas[k], ad = pre[k - 1][3];
        const double invA = 1.0 / pre[k][3];
        dep = invA;
        const double pA_s = l.X * ad + l.Y * (1.0 - ad) - l.Z * ad;
        const double pA_d = l.X * as - l.Y * as + l.Z * (1.0 - as);
        double D[3], gs[3], gd[3];
        blend3_all(mode, cs, cd, D, gs, gd);
        // the colour model of layer k: the colour part of dE,
        // alpha_k 2 P (c_k - mu_k), and the distance for the energy
        double gE[3];
        distance_gradient(cs[0], cs[1], cs[2], l, gE);
        dist[k] = distance(cs[0], cs[1], cs[2], l);
        double adj_as = 0.0, adj_ad = 0.0;
        double nB[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double B = pre[k][c];
            const double dB_dcs = (as * ad * gs[c] + l.Y * (1.0 - ad) * as) * invA;
            const double dB_das = ((D[c] * ad + l.Y * (1.0 - ad) * cs[c] - l.Z * ad * cd[c]) - B * pA_s) * invA;
            const double dB_dcd = (as * ad * gd[c] + l.Z * (1.0 - as) * ad) * invA;
            const double dB_dad = ((D[c] * as - l.Y * as * cs[c] + l.Z * (1.0 - as) * cd[c]) - B * pA_d) * invA;
            grad[L + 3 * k + c] = alphas[k] * gE[c] + dB_dcs * aB[c];
            adj_as += dB_das * aB[c];
            adj_ad += dB_dad * aB[c];
            nB[c] = dB_dcd * aB[c];
        }
        double ga = adj_as + pA_s * aA;
        if constexpr (REFINE) ga = ga + (rho * (alphas[k] - ta[k]) - lam[3 + k]);
        grad[k] = (REFINE ? dist[k] : 0.01) + ga;
        aprev = aA;
        aA = adj_ad + pA_d * aA;
        aB[0] = nB[0];
        aB[1] = nB[1];
        aB[2] = nB[2];
    }
    double ga0 = aA;
    if constexpr (REFINE) ga0 = ga0 + (rho * (alphas[0] - ta[0]) - lam[3]);
    {
        const DevLayer& l0 = layer_after(lay, 0, dep);
        dist[0] = distance(colors[0], colors[1], colors[2], l0);
        grad[0] = (REFINE ? dist[0] : 0.01) + ga0;
        double gE[3];
        distance_gradient(colors[0], colors[1], colors[2], l0, gE);
#pragma unroll
        for (int c = 0; c < 3; ++c) grad[L + c] = alphas[0] * gE[c] + aB[c];
    }
    if constexpr (GRAY) {
        const double* c = colors + 3;
        const double norm = __builtin_sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        if (norm > 1e-3) {
            const double inv = 1.0 / norm;
            double wg[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) wg[j] = rho * C[NB + j] - lam[NB + j];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double t = 0.0;
#pragma unroll
                for (int j = 0; j < 3; ++j) t = t + (kSqrt3 * (i == j ? 1.0 : 0.0) - inv * c[i]) * wg[j];
                grad[L + 3 + i] = grad[L + 3 + i] + t;
            }
        }
    }
    // the energy, its terms summed bottom to top as the reference does
    double energy = 0.0;
#pragma unroll
    for (int k = 0; k < L; ++k) energy += alphas[k] * dist[k];
    if constexpr (SPARSE && !REFINE) {
        double sum = 0.0, squared = 0.0;
#pragma unroll
        for (int i = 0; i < L; ++i) sum += alphas[i];
#pragma unroll
        for (int i = 0; i < L; ++i) squared += alphas[i] * alphas[i];
        energy += sparse_sigma * ((sum / squared) - 1.0);
    }
    if constexpr (!REFINE) {
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < L; ++i) sum += alphas[i];
        energy += 0.01 * sum;
    }
    return energy + (-lagrange) + 0.5 * rho * squared;
}

}  // namespace ubhip
