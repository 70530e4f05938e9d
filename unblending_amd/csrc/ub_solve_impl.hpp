// ub_solve_impl.hpp -- the per-pixel solve kernels on gfx950, one pixel per lane
// (templates shared by ub_solve.hip, which launches the plain problem, and
// ub_solve_var.hip, which launches the reference's problem variants).
//
// Each lane runs, for its pixel, the reference's augmented-Lagrangian loop
// (unblending/src/unblending.cpp:155-309: rho0 = 100, beta = 10, gamma = 0.25,
// eps = 5e-3, at most 22 outer iterations, the last iterate returned) around
// the project's projected L-BFGS inner minimiser (the stand-in for NLopt
// LD_LBFGS at unblending.cpp:249: maxeval 1000, ftol_rel = xtol_rel = 5e-3):
//   free set -> two-loop recursion -> projected Armijo backtracking with
//   quadratic interpolation in [0.1 t, 0.5 t] -> (s, y) update with a
//   curvature test -> NLopt's relative ftol / xtol tests.
// The layer set is wave-uniform and read with scalar loads (the simple kernel
// from its argument segment, the persistent kernel from the image descriptor
// through the constant address space).
// The one-pair minimiser keeps its (s, y) pair in registers; longer histories
// (the default: NLopt's default vector storage keeps every pair of an inner
// solve) live in per-lane rings in the workspace, and the wave forms the
// search directions of its lanes cooperatively, 8 lanes a pixel (see
// direction_service).  The two-loop recursion's dot products are summed in the
// order orc_tree_dot of the oracle defines: eight partial sums of four
// consecutive components, then pairwise.
#pragma once

#include "ub_kernels.hpp"

namespace ubhip {
namespace {

__device__ __forceinline__ double clampd(double v, double lo, double hi)
{
    v = (v < lo) ? lo : v;
    return (v > hi) ? hi : v;
}

__device__ __forceinline__ bool relstop(double vold, double vnew, double reltol)
{
    return __builtin_fabs(vnew - vold) < reltol * (__builtin_fabs(vnew) + __builtin_fabs(vold)) * 0.5 ||
           vnew == vold;
}

// The two-loop recursion's dot product: partial sums of components
// 4j .. 4j + 3 (j = 0..7), each in index order from 0.0, added pairwise.
template <int N>
__device__ __forceinline__ double tree_dot(const double* a, const double* b)
{
    static_assert(N <= 32, "four components a partial sum");
    double p[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (4 * j + c < N) s = s + a[4 * j + c] * b[4 * j + c];
        }
        p[j] = s;
    }
    return ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
}

// Per-variable bounds (unblending.cpp:174-208): [0, 1]; alpha_0 = 1 when
// opaque; alphas at their targets when fixed; background colour fixed when
// smoothed.
template <int L, bool REFINE>
struct Bounds {
    bool opaque, smooth, alpha_fixed;
    const double* ta;
    const double* bg;
    __device__ __forceinline__ void get(int i, double& lo, double& hi) const
    {
        lo = 0.0;
        hi = 1.0;
        if (i < L) {
            if (REFINE && alpha_fixed) { lo = ta[i]; hi = ta[i]; }
            if (i == 0 && opaque) {
                lo = 1.0;
                if (REFINE && alpha_fixed) hi = 1.0;
            }
        } else if (REFINE && smooth && i < L + 3) {
            lo = bg[i - L];
            hi = bg[i - L];
        }
    }
};

template <int L, bool REFINE>
struct Problem {
    const DevLayer* lay;
    double tgt[3];
    double ta[L];
    double lam[Shape<L, REFINE>::NC];
    double rho;
    __device__ __forceinline__ double eval(const double* x, double* g) const
    {
        double cc[4];
        return objective<L, REFINE>(x, lay, tgt, ta, lam, rho, g, cc);
    }
};

// The inner minimiser (the host oracle's orc_minimize, same arithmetic).
template <int L, bool REFINE, int M>
__device__ void minimize(double* x, const Bounds<L, REFINE>& bd, const Problem<L, REFINE>& pb,
                         const SolverParams& sp, int& nev_out, int& iters_out)
{
    constexpr int N = 4 * L;
    double g[N], d[N], q[N], xt[N], gt[N];
    double S[M][N], Yh[M][N], rh[M], a[M];
    bool freev[N];
    int count = 0;
    double gamma = 1.0;

#pragma unroll
    for (int i = 0; i < N; ++i) {
        double lo, hi;
        bd.get(i, lo, hi);
        x[i] = clampd(x[i], lo, hi);
    }
    double f = pb.eval(x, g);
    int nev = 1, iters = 0;

    for (;;) {
        bool finite = __builtin_isfinite(f);
#pragma unroll
        for (int i = 0; i < N; ++i) finite = finite && __builtin_isfinite(g[i]);
        if (!finite) break;

        double pgmax = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double lo, hi;
            bd.get(i, lo, hi);
            freev[i] = lo < hi && !(x[i] <= lo && g[i] > 0.0) && !(x[i] >= hi && g[i] < 0.0);
            q[i] = freev[i] ? g[i] : 0.0;
            const double m = __builtin_fabs(q[i]);
            pgmax = (m > pgmax) ? m : pgmax;
        }
        if (pgmax == 0.0) break;
        ++iters;

#pragma unroll
        for (int k = M - 1; k >= 0; --k) {
            if (M - 1 - k < count) {
                a[k] = rh[k] * tree_dot<N>(S[k], q);
#pragma unroll
                for (int i = 0; i < N; ++i) q[i] = q[i] - a[k] * Yh[k][i];
            }
        }
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = gamma * q[i];
#pragma unroll
        for (int k = 0; k < M; ++k) {
            if (k >= M - count) {
                const double b = rh[k] * tree_dot<N>(Yh[k], d);
#pragma unroll
                for (int i = 0; i < N; ++i) d[i] = d[i] + (a[k] - b) * S[k][i];
            }
        }
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = freev[i] ? -d[i] : 0.0;
        double gd = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) gd += g[i] * d[i];
        if (!(gd < 0.0)) {
            count = 0;
            gamma = 1.0;
#pragma unroll
            for (int i = 0; i < N; ++i) d[i] = freev[i] ? -g[i] : 0.0;
        }

        double t = 1.0;
        if (count == 0) {
            double dd = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) dd += d[i] * d[i];
            const double inv = 1.0 / __builtin_sqrt(dd);
            t = (inv < 1.0) ? inv : 1.0;
        }

        bool accepted = false;
        double ft = 0.0;
        for (int ls = 0;; ++ls) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double lo, hi;
                bd.get(i, lo, hi);
                xt[i] = clampd(x[i] + t * d[i], lo, hi);
            }
            ft = pb.eval(xt, gt);
            ++nev;
            double dphi = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) dphi += g[i] * (xt[i] - x[i]);
            if (__builtin_isfinite(ft) && ft <= f + 1e-4 * dphi) {
                accepted = true;
                break;
            }
            if (nev >= sp.max_evals || ls + 1 >= sp.max_ls) break;
            double tn;
            if (dphi < 0.0) {
                const double denom = 2.0 * (ft - f - dphi);
                tn = (-dphi * t) / denom;
            } else {
                tn = 0.5 * t;
            }
            const double lo = 0.1 * t, hi = 0.5 * t;
            tn = (tn >= lo) ? tn : lo;
            tn = (tn > hi) ? hi : tn;
            t = tn;
        }
        if (!accepted) break;

        double sy = 0.0, yy = 0.0;
        bool xstop = true;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double s = xt[i] - x[i];
            const double y = gt[i] - g[i];
            sy += s * y;
            yy += y * y;
            xstop = xstop && (__builtin_fabs(s) < sp.xtol * (__builtin_fabs(xt[i]) + __builtin_fabs(x[i])) * 0.5 ||
                              xt[i] == x[i]);
            x[i] = xt[i];
            g[i] = gt[i];
            xt[i] = s;  // keep s and y for the history update
            gt[i] = y;
        }
        const double fold = f;
        f = ft;
        if (sy > 2.220446049250313e-16 * yy) {
#pragma unroll
            for (int k = 0; k + 1 < M; ++k) {
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    S[k][i] = S[k + 1][i];
                    Yh[k][i] = Yh[k + 1][i];
                }
                rh[k] = rh[k + 1];
            }
#pragma unroll
            for (int i = 0; i < N; ++i) {
                S[M - 1][i] = xt[i];
                Yh[M - 1][i] = gt[i];
            }
            rh[M - 1] = 1.0 / sy;
            gamma = sy / yy;
            count = (count < M) ? count + 1 : M;
        } else {
            count = 0;
            gamma = 1.0;
        }
        if (nev >= sp.max_evals) break;
        if (relstop(fold, f, sp.ftol)) break;
        if (xstop) break;
    }
    nev_out = nev;
    iters_out = iters;
}

// Constraint vector alone (equations.cpp:113-155), for the outer loop;
// GRAY: layer 1's gray rows after the alpha rows.
template <int L, bool REFINE, bool GRAY = false, class TG, class TA>
__device__ __forceinline__ void constraints(const double* x, const DevLayer* lay, const TG& tgt, const TA& ta,
                                            double* C)
{
    const double* alphas = x;
    const double* colors = x + L;
    double acc0 = colors[0], acc1 = colors[1], acc2 = colors[2], accA = alphas[0];
#pragma unroll
    for (int k = 1; k < L; ++k) {
        const DevLayer& l = lay[k];
        const int mode = uniform_mode(l.mode);
        const double as = alphas[k], ad = accA;
        const double a = l.X * as * ad + l.Y * as * (1.0 - ad) + l.Z * ad * (1.0 - as);
        const double ky = l.Y * as * (1.0 - ad);
        const double kz = l.Z * ad * (1.0 - as);
        double cd[3] = {acc0, acc1, acc2};
        double out[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double cs = colors[3 * k + c];
            const double f = blend(cs, cd[c], mode);
            const double p = f * as * ad + ky * cs + kz * cd[c];
            out[c] = (a > 1e-12) ? p / a : p;
        }
        acc0 = out[0];
        acc1 = out[1];
        acc2 = out[2];
        accA = a;
    }
    C[0] = acc0 - tgt[0];
    C[1] = acc1 - tgt[1];
    C[2] = acc2 - tgt[2];
    if constexpr (REFINE) {
#pragma unroll
        for (int i = 0; i < L; ++i) C[3 + i] = alphas[i] - ta[i];
    } else {
        C[3] = accA - 1.0;
    }
    if constexpr (GRAY) gray_rows(colors + 3, C + Shape<L, REFINE, GRAY>::NB);
}

template <int NC>
__device__ __forceinline__ double norm(const double* v)
{
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NC; ++j) s += v[j] * v[j];
    return __builtin_sqrt(s);
}

// Per-pixel inputs and the initial solution (unblending.cpp:185-208, 401-428):
// target colour; refinement targets = crop + normalisation of the filtered
// alphas (unblending.cpp:366-389, 311-333) and the cropped smoothed
// background; x = default or refinement initialisation.
template <int L, bool REFINE, class Src>
__device__ __forceinline__ void load_pixel(const Src& args, const DevLayer* lay, long long i, double tgt[3],
                                           double ta[L], double bg[3], double* x, bool use_warm = true)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) tgt[c] = args.rgb[3 * i + c];
    bg[0] = bg[1] = bg[2] = 0.0;
    if constexpr (REFINE) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
            double v = args.filtered[(long long)l * args.n + i];
            ta[l] = (v < 0.0) ? 0.0 : ((v > 1.0) ? 1.0 : v);
        }
        if (args.all_so && __builtin_fabs(ta[0] - 1.0) < 1e-05) {
        } else if (args.all_plus) {
            double sum = 0.0;
#pragma unroll
            for (int l = 0; l < L; ++l) sum += ta[l];
#pragma unroll
            for (int l = 0; l < L; ++l) ta[l] = ta[l] / sum;
        } else {
#pragma unroll
            for (int l = 0; l < L; ++l) ta[l] = 0.0;
        }
        if (args.smooth) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double v = args.filtered[(long long)(L + c) * args.n + i];
                bg[c] = (v < 0.0) ? 0.0 : ((v > 1.0) ? 1.0 : v);
            }
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            x[l] = ta[l];
            const double* px = args.first + 4 * ((long long)l * args.n + i);
#pragma unroll
            for (int c = 0; c < 3; ++c) x[L + 3 * l + c] = px[c];
        }
    } else if (args.warm && use_warm) {
        // warm start: the predicted first pass
#pragma unroll
        for (int l = 0; l < L; ++l) {
            ta[l] = 0.0;
            const double* px = args.warm + 4 * ((long long)l * args.n + i);
            x[l] = px[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) x[L + 3 * l + c] = px[c];
        }
    } else {
#pragma unroll
        for (int l = 0; l < L; ++l) {
            ta[l] = 0.0;
            x[l] = 0.50;
#pragma unroll
            for (int c = 0; c < 3; ++c) x[L + 3 * l + c] = lay[l].mean[c];
        }
    }
    if (args.opaque) x[0] = 1.0;
    if (REFINE && args.smooth) {
#pragma unroll
        for (int c = 0; c < 3; ++c) x[L + c] = bg[c];
    }
}

template <int L, bool REFINE, int M>
__global__ void __launch_bounds__(64) solve_simple_kernel(PixelArgs args)
{
    constexpr int N = 4 * L;
    constexpr int NC = Shape<L, REFINE>::NC;
    const DevLayer* lay = args.layers;  // uniform: scalar loads from the argument segment

    const long long pix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = pix < args.n;
    const long long i = active ? pix : 0;

    Problem<L, REFINE> pb;
    pb.lay = lay;
    double bgc[3];
    double x[N];
    load_pixel<L, REFINE>(args, lay, i, pb.tgt, pb.ta, bgc, x);

    Bounds<L, REFINE> bd;
    bd.opaque = args.opaque != 0;
    bd.smooth = args.smooth != 0;
    bd.alpha_fixed = args.sp.alpha_fixed != 0;
    bd.ta = pb.ta;
    bd.bg = bgc;

#pragma unroll
    for (int j = 0; j < NC; ++j) pb.lam[j] = 0.0;
    pb.rho = args.sp.rho0;

    unsigned long long evals = 0, iters = 0, outer = 0;
    if (active) {
        int count = 0;
        for (;;) {
            double xn[N];
#pragma unroll
            for (int v = 0; v < N; ++v) xn[v] = x[v];
            int ev = 0, it = 0;
            minimize<L, REFINE, M>(xn, bd, pb, args.sp, ev, it);
            evals += ev;
            iters += it;
            outer += 1;
            double g[NC], gn[NC];
            constraints<L, REFINE>(x, lay, pb.tgt, pb.ta, g);
            constraints<L, REFINE>(xn, lay, pb.tgt, pb.ta, gn);
#pragma unroll
            for (int j = 0; j < NC; ++j) pb.lam[j] = pb.lam[j] - pb.rho * gn[j];
            const double ngn = norm<NC>(gn);
            if (ngn > args.sp.gamma * norm<NC>(g)) pb.rho = pb.rho * args.sp.beta;
            double dx[N];
#pragma unroll
            for (int v = 0; v < N; ++v) dx[v] = xn[v] - x[v];
            const bool unchanged = norm<N>(dx) < args.sp.eps;
            const bool satisfied = ngn < args.sp.eps;
#pragma unroll
            for (int v = 0; v < N; ++v) x[v] = xn[v];
            if ((unchanged && satisfied) || count > args.sp.max_count) break;
            ++count;
        }
#pragma unroll
        for (int l = 0; l < L; ++l) {
            double* px = args.out + 4 * ((long long)l * args.n + pix);
            px[0] = x[L + 3 * l];
            px[1] = x[L + 3 * l + 1];
            px[2] = x[L + 3 * l + 2];
            px[3] = x[l];
            if (args.alpha_planes) args.alpha_planes[(long long)l * args.n + pix] = x[l];
        }
    }
    if (args.counters) {
        // wave-level reduction, one atomic per counter per wave
        for (int off = 32; off > 0; off >>= 1) {
            evals += __shfl_xor(evals, off);
            iters += __shfl_xor(iters, off);
            outer += __shfl_xor(outer, off);
        }
        const unsigned long long pixels = __popcll(__ballot(active));
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&args.counters[kEvals], evals);
            atomicAdd(&args.counters[kIterations], iters);
            atomicAdd(&args.counters[kOuter], outer);
            atomicAdd(&args.counters[kPixels], pixels);
        }
    }
}

// ---------------------------------------------------------------------------
// Persistent, evaluation-synchronous solve over a batch of images.
//
// Every trip around the kernel's loop evaluates the objective once for every
// lane; between evaluations each lane advances its own pixel's state machine
// (start of an inner solve, line-search trial, search direction, outer
// update, next pixel).  A lane whose pixel is finished takes the next item of
// the wave's current chunk, and the wave takes the next chunk from a device
// counter, so lanes stay busy however uneven the per-pixel iteration counts
// are.  A wave solves one image at a time (the layer table is wave-uniform):
// when its next chunk belongs to another image, lanes that finish wait until
// the wave has drained, and the wave then moves on.  A launch holds images
// of one layer count (the kernel is compiled per layer count, which keeps
// each one's register allocation its own); one image is a batch of one.  The arithmetic
// is the same sequence, per pixel, as solve_simple_kernel and the host
// oracle.
//
// Registers hold the current point and the (s, y) pair; the pair's registers
// hold the trial point and trial gradient from the projection that starts a
// trip until the trial is accepted (the pair is no longer needed once the
// direction is computed).  LDS holds, per lane, the gradient, the search
// direction, the multipliers, the constraints that depend on the composite
// at the current point (kept from that point's evaluation) and, in
// refinement where the layer count leaves room, the pixel's target colour and
// target alphas, so that an evaluation's live set stays near the 256
// architectural registers of a lane (the excess goes to accumulation
// registers, not to scratch).
// Only the outer iterate, read and written once per outer iteration, has a
// per-lane global slot.
// ---------------------------------------------------------------------------
// The reference's problem variants a persistent kernel is compiled for (bit
// flags): layer 1 gray, the first pass's sparsity term.
constexpr int kVarGray = 1, kVarSparse = 2;

// per-lane phases of the persistent solve (distinct from the Counter slots)
enum Phase : int {
    kPhNeedPixel = 0, kPhStart = 1, kPhTrial = 2, kPhWait = 3, kPhOuter = 4, kPhRetry = 5,
    kPhDir = 6  // history: search direction in progress (direction_service)
};

// Lanes waiting for a pixel, or for their outer update, are served together
// once at least this many wait (or nothing else is left to do).
constexpr int kFetchBatch = 8;
constexpr int kOuterBatch = 8;

// A per-lane vector in LDS, laid out [element][64 lanes].
struct LaneVec {
    double* p;
    __device__ __forceinline__ double& operator[](int i) const { return p[i * 64]; }
};

// The pixel's constant inputs read by every evaluation, the target colour and
// the target alphas (refinement), in LDS or in registers.
template <int L, bool IN_LDS>
struct PixelConst;

template <int L>
struct PixelConst<L, true> {
    static constexpr int kCount = 3 + L;
    LaneVec tgt, ta;
    __device__ __forceinline__ PixelConst(double* base) : tgt{base}, ta{base + 3 * 64} {}
};

template <int L>
struct PixelConst<L, false> {
    static constexpr int kCount = 1;
    double tgt[3], ta[L];
    __device__ __forceinline__ PixelConst(double*) {}
};

template <int L, bool REFINE, int VAR = 0>
struct PersistentShape {
    static constexpr int N = 4 * L;
    static constexpr int NC = Shape<L, REFINE, (VAR & kVarGray) != 0>::NC;
    static constexpr int NCC = REFINE ? 3 : 4;  // constraints that need the composite
    // doubles of LDS a lane may use with four workgroups of 64 lanes on a CU
    static constexpr int kLaneBudget = 160 * 1024 / (4 * 64 * 8);
    // the refinement's larger live set leaves no registers for the pixel's
    // targets; the first pass keeps them in registers
    static constexpr bool kPixelInLds = REFINE && 2 * N + NC + NCC + (3 + L) <= kLaneBudget;
};

template <int L, bool REFINE, int VAR = 0>
constexpr int persistent_lds_doubles()
{
    using PS = PersistentShape<L, REFINE, VAR>;
    return 2 * PS::N + PS::NC + PS::NCC + PixelConst<L, PS::kPixelInLds>::kCount;
}

static_assert(persistent_lds_doubles<7, true>() <= PersistentShape<7, true>::kLaneBudget, "LDS budget");
static_assert(persistent_lds_doubles<8, true>() <= PersistentShape<8, true>::kLaneBudget, "LDS budget");
static_assert(persistent_lds_doubles<8, false>() <= PersistentShape<8, false>::kLaneBudget, "LDS budget");

__device__ __forceinline__ long long uniform64(unsigned long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

typedef const __attribute__((address_space(4))) ImageDesc ImageDescC;

__device__ __forceinline__ ImageDescC* as_constant(const ImageDesc* p) { return (ImageDescC*)p; }

// Wave-uniform position in the launch's chunk queue.
struct WaveQueue {
    int img;                // image being solved; -1 once the queue is empty
    long long next, end;    // items of the current chunk not yet handed out
    int pimg;               // a claimed chunk of another image, started once
    long long pnext, pend;  // the wave has drained
    bool pending, exhausted;
};

// A lane's work counts (32 bits: a lane solves far fewer than 2^32 /
// 1000 pixels of a launch) and the wave's trips around the evaluation loop
// (wave-uniform).
struct LaneCounters {
    unsigned evals, iters, outer, pixels, pairs;
    unsigned long long trips;
};

// Claims the next non-empty chunk of the queue: its image and item range.
// Returns false when the queue is exhausted.  Wave-uniform.
__device__ __forceinline__ bool claim_chunk(const BatchArgs& a, int& img, long long& b, long long& e)
{
    for (;;) {
        unsigned long long k = 0;
        if (threadIdx.x == 0) k = atomicAdd(&a.counters[a.queue], 1ull);
        const long long kk = uniform64(__shfl(k, 0));
        if (kk >= a.chunk_count) return false;
        int i = a.image_count - 1;
        ImageDescC* images = as_constant(a.images);
        while (i > 0 && images[i].chunk_base > kk) --i;
        ImageDescC* d = images + i;
        const long long lim = d->list ? (long long)*d->list_count : d->n;
        const long long bb = (kk - d->chunk_base) * (long long)a.chunk_size;
        const long long ee = (bb + a.chunk_size < lim) ? bb + a.chunk_size : lim;
        if (bb < ee) {
            img = i;
            b = bb;
            e = ee;
            return true;
        }
    }
}

// A double moved between lanes by a DPP pattern (both halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over a 4-lane group (a quad) of the lanes' values A_l, where lane l's
// A_l = p_2l + p_2l+1: quad_perm [1,0,3,2] then [2,3,0,1], giving
// ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)) -- the tree order of the
// oracle's orc_tree_dot.  Every lane of the quad ends with the same value
// (each addition's operands are the same pair, in either order).
__device__ __forceinline__ double quad_sum(double v)
{
    v = v + dpp_f64<0xB1>(v);
    v = v + dpp_f64<0x4E>(v);
    return v;
}

// Eight zeros: what a lane reads for components past the end of s or y.
__device__ double kZeros[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};

// Records of one history ring streamed per load batch of the two-loop
// recursion (three at L = 8, whose refinement kernel already spills), and the
// recursion steps a pending direction advances per trip.
constexpr int kHistBlock(int L) { return L >= 8 ? 3 : 4; }
constexpr int kDirBudget = 48;

// Search directions from the (s, y) history for the lanes in `pend` (those
// whose direction is pending).  A direction of a pixel with `count` records
// is 2 count steps of the two-loop recursion -- steps 0 .. count - 1 newest
// to oldest (a_k = rho_k s_k.q, q -= a_k y_k), then q = gamma q, steps
// count .. 2 count - 1 oldest to newest (b = rho_k y_k.q, q += (a_k - b) s_k);
// `dstep` counts the steps done, and the vector being transformed stays in
// the pixel's LDS column dv[i * 64 + lane] between trips.  Each call advances
// every pending direction by at most kDirBudget steps, so that a pixel with a
// long history does not hold up the evaluations of the rest of the wave.
// The wave's 16 quads serve the pending pixels, longest remaining work first,
// and a quad whose pixel is done takes the next one at once (so that a long
// direction does not leave the other quads idle): lane l of quad g (lanes
// 4g .. 4g + 3) holds components 8l .. 8l + 7 of the quad's pixel and
// streams that pixel's records -- the k-th oldest at
// slot (head + k) mod cap of its ring hist + p * cap * R -- kHistBlock
// records a load batch; the batch that ends the first loop also runs the
// second loop's first steps on the records it holds (the oldest), their
// multipliers kept in registers.  4L is a multiple of 4, so each half (four
// components) of a lane's share is wholly inside the vector or wholly past
// its end; a half past the end reads zeros (kZeros) and takes no part in the
// arithmetic.  A dot product
// is each lane's two four-component partial sums, added, then summed over
// the quad (quad_sum).  A quad's trip counts are uniform over the quad, so
// its DPP moves run on whole quads.  Returns the lanes whose direction is
// complete.
template <int L>
__device__ __forceinline__ unsigned long long direction_service(unsigned long long pend, double* dvcol,
                                                                const double* hblock, int cap, int count,
                                                                int head, double gamma, int& dstep)
{
    constexpr int N = 4 * L;
    constexpr int R = kHistRecord(L);
    constexpr int U = kHistBlock(L);
    const int lane = threadIdx.x, l = lane & 3;
    const bool live = 8 * l < N;       // components 8l .. 8l + 3 exist
    const bool upper = 8 * l + 8 <= N;  // components 8l + 4 .. 8l + 7 exist
    const bool mine = (pend >> lane) & 1ull;
    const int npend = __popcll(pend);
    // rank among the pending lanes: most remaining steps first (the other
    // lanes rank after them, in lane order, so that the ranks are a
    // permutation of the lanes)
    const int rem = 2 * count - dstep;
    int prank = 0;
    {
        unsigned long long m = pend;
        while (m) {
            const int o = __builtin_ctzll(m);
            m &= m - 1ull;
            const int ro = __builtin_amdgcn_readlane(rem, o);
            prank += (ro > rem || (ro == rem && o < lane)) ? 1 : 0;
        }
    }
    const int rank = mine ? prank : npend + __popcll(~pend & ((1ull << lane) - 1ull));
    // lane k of `order` holds the lane of rank k
    const int order = __builtin_amdgcn_ds_permute(rank << 2, lane);
    // lane k: the state of the rank-k pixel
    const int o_cnt = __shfl(count, order);
    const int o_hd = __shfl(head, order);
    const int o_ds = __shfl(dstep, order);
    const double o_gam = __shfl(gamma, order);
    // the records written at this trip's acceptances, by other lanes
    __threadfence_block();
    // the pixel a quad serves (the lanes of a quad agree on all of these);
    // a quad whose pixel is done for this call takes the next one by rank
    bool has = false;
    int p = 0, cnt = 0, hd = 0, pos = 0, end = 0;
    double gam = 1.0;
    double* ring = const_cast<double*>(hblock);
    double qv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) qv[c] = 0.0;
    int next = 0;  // the next rank handed out (wave-uniform)
    constexpr unsigned long long kQuadLead = 0x1111111111111111ull;
    for (;;) {
        const bool idle = !(pos < end);
        unsigned long long fin = __ballot(idle && has);
        if (fin) {
            // the finished pixels' vectors back to their LDS columns, their
            // step counts to their lanes
            if (idle && has) {
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    if (8 * l + c < N) dvcol[(8 * l + c) * 64 + p] = qv[c];
                }
            }
            fin &= kQuadLead;
            while (fin) {
                const int o = __builtin_ctzll(fin);
                fin &= fin - 1ull;
                const int po = __builtin_amdgcn_readlane(p, o);
                const int so = __builtin_amdgcn_readlane(pos, o);
                if (lane == po) dstep = so;
            }
            if (idle) has = false;
        }
        const unsigned long long qm = __ballot(idle) & kQuadLead;
        if (next < npend && qm) {
            const int nr = next + __popcll(qm & ((1ull << (lane & ~3)) - 1ull));
            const int nl = nr < 64 ? nr : 0;
            const int src = __shfl(order, nl);
            const int pc = __shfl(o_cnt, nl);
            const int ph = __shfl(o_hd, nl);
            const int pd = __shfl(o_ds, nl);
            const double pg = __shfl(o_gam, nl);
            next += __popcll(qm);
            if (idle && nr < npend) {
                has = true;
                p = src;
                cnt = pc;
                hd = ph;
                gam = pg;
                pos = pd;
                end = min(2 * cnt, pos + kDirBudget);
                ring = const_cast<double*>(hblock) + (long long)p * cap * R;
#pragma unroll
                for (int c = 0; c < 8; ++c) qv[c] = (8 * l + c < N) ? dvcol[(8 * l + c) * 64 + p] : 0.0;
                if (cnt == 0) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) qv[c] = gam * qv[c];
                }
            }
        }
        if (!__ballot(pos < end)) break;
        // a batch stays within one loop of the recursion (its loads of a_k
        // follow the stores of the first loop in the wave's program order,
        // which the wave's memory operations to one address keep)
        const int lim = (pos < cnt && cnt < end) ? cnt : end;
        const int take = min(U, lim - pos);  // <= 0 for a quad whose work is done
        const bool back = pos >= cnt;        // second loop, oldest to newest
        // Every lane loads U records a batch: record indices are clamped
        // to the ring (a quad with fewer records left re-reads one), so the
        // loads are unconditional and their registers carry nothing from
        // one batch to the next; records past `take` are not computed.
        // The dot-product vector is s in the first loop and y in the
        // second, the update vector the other one; a half of a lane's share
        // past the vector's end reads zeros (kZeros) and is not computed.
        double* recp[U];
        double rv[U], av[U], dA[U][8], uB[U][8];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int k = back ? pos + u - cnt : cnt - 1 - (pos + u);
            k = (k < 0) ? 0 : k;
            k = (k >= cnt) ? ((cnt > 0) ? cnt - 1 : 0) : k;
            int slot = hd + k;
            if (slot >= cap) slot -= cap;
            double* const rec = ring + (long long)slot * R;
            recp[u] = rec;
            rv[u] = rec[2 * N];
            av[u] = rec[2 * N + 1];
            const double* const pa = rec + (back ? N : 0) + 8 * l;
            const double* const pb = rec + (back ? 0 : N) + 8 * l;
            const double2* const a_lo = reinterpret_cast<const double2*>(live ? pa : kZeros);
            const double2* const a_hi = reinterpret_cast<const double2*>(upper ? pa + 4 : kZeros);
            const double2* const b_lo = reinterpret_cast<const double2*>(live ? pb : kZeros);
            const double2* const b_hi = reinterpret_cast<const double2*>(upper ? pb + 4 : kZeros);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const double2 x0 = a_lo[h], x1 = a_hi[h], z0 = b_lo[h], z1 = b_hi[h];
                dA[u][2 * h] = x0.x;
                dA[u][2 * h + 1] = x0.y;
                dA[u][4 + 2 * h] = x1.x;
                dA[u][4 + 2 * h + 1] = x1.y;
                uB[u][2 * h] = z0.x;
                uB[u][2 * h + 1] = z0.y;
                uB[u][4 + 2 * h] = z1.x;
                uB[u][4 + 2 * h + 1] = z1.y;
            }
        }
        if (back && take > 0 && pos == cnt) {
#pragma unroll
            for (int c = 0; c < 8; ++c) qv[c] = gam * qv[c];
        }
        // the batch that ends the first loop (its records are the oldest
        // `take`) also runs the second loop's first `take` steps on them, the
        // multipliers kept in registers, when the trip's budget has room
        const bool turn = !back && take > 0 && pos + take == cnt && pos + 2 * take <= end;
        double ak[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ak[u] = 0.0;
            if (u < take) {
                // a half past the vector's end takes no part: its partial
                // sum is 0.0 and its components of q stay 0.0, as in the
                // oracle's tree order
                double p0 = 0.0, p1 = 0.0;
                if (live) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) p0 = p0 + dA[u][c] * qv[c];
                }
                if (upper) {
#pragma unroll
                    for (int c = 4; c < 8; ++c) p1 = p1 + dA[u][c] * qv[c];
                }
                const double dot = quad_sum(p0 + p1);
                const double sc = rv[u] * dot;  // a_k (first loop) or b (second)
                double coef;
                if (!back) {
                    if (l == 0 && !turn) recp[u][2 * N + 1] = sc;
                    ak[u] = sc;
                    coef = -sc;  // q - a y, as q + (-a) y: the same rounding
                } else {
                    coef = av[u] - sc;
                }
                if (live) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) qv[c] = qv[c] + coef * uB[u][c];
                }
                if (upper) {
#pragma unroll
                    for (int c = 4; c < 8; ++c) qv[c] = qv[c] + coef * uB[u][c];
                }
            }
        }
        if (turn) {
#pragma unroll
            for (int c = 0; c < 8; ++c) qv[c] = gam * qv[c];
            // oldest first: record k = take - 1 - u sits in slot u
#pragma unroll
            for (int u = U - 1; u >= 0; --u) {
                if (u < take) {
                    double p0 = 0.0, p1 = 0.0;
                    if (live) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) p0 = p0 + uB[u][c] * qv[c];
                    }
                    if (upper) {
#pragma unroll
                        for (int c = 4; c < 8; ++c) p1 = p1 + uB[u][c] * qv[c];
                    }
                    const double dot = quad_sum(p0 + p1);
                    const double coef = ak[u] - rv[u] * dot;
                    if (live) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) qv[c] = qv[c] + coef * dA[u][c];
                    }
                    if (upper) {
#pragma unroll
                        for (int c = 4; c < 8; ++c) qv[c] = qv[c] + coef * dA[u][c];
                    }
                }
            }
        }
        pos += (take > 0) ? (turn ? 2 * take : take) : 0;
    }
    return __ballot(mine && dstep == 2 * count);
}

// The solve of the wave's current image and of the images that follow it in
// the queue (all with L layers).  Returns when the queue is exhausted.
template <int L, bool REFINE, bool HIST, int VAR>
__device__ __forceinline__ void run_images(const BatchArgs& args, double* smem, WaveQueue& q, LaneCounters& lc)
{
    constexpr bool GRAY = (VAR & kVarGray) != 0;
    constexpr bool SPARSE = !REFINE && (VAR & kVarSparse) != 0;
    using PS = PersistentShape<L, REFINE, VAR>;
    constexpr int N = PS::N;
    constexpr int NC = PS::NC;
    using PC = PixelConst<L, PS::kPixelInLds>;
    const int lane = threadIdx.x;
    // per-lane LDS vectors: gradient at the current point, search direction,
    // multipliers, composite constraints at the current point, pixel targets
    const LaneVec gv{smem + lane};
    const LaneVec dv{smem + N * 64 + lane};
    const LaneVec lam{smem + 2 * N * 64 + lane};
    const LaneVec cc{smem + (2 * N + NC) * 64 + lane};
    PC pc(smem + (2 * N + NC + PS::NCC) * 64 + lane);
    double bg[3] = {0.0, 0.0, 0.0};  // background colour (smoothed refinement)

    const bool afixed = REFINE && args.sp.alpha_fixed != 0;
    bool opaque = false, smooth = false;

    // S, Y: the (s, y) pair, and between direction and acceptance the trial
    // point and trial gradient.
    double xc[N], S[N], Y[N];
    double rh = 0.0;
    double f = 0.0, t = 1.0, rho = 0.0, gamma = 1.0, cn_outer = 0.0;
    int phase = kPhNeedPixel, count = 0, nev = 0, ls = 0, ocount = 0;
    int head = 0;  // history: slot of the oldest record of the ring
    int dstep = 0;  // history: two-loop steps done of the pending direction
    unsigned int freev = 0;  // free variables of the direction being formed
    bool warm_try = false;  // this pixel's solve started from the warm-start prediction
    long long pix = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        xc[i] = 0.0;
        gv[i] = 0.0;
        S[i] = 0.0;
        Y[i] = 0.0;
    }
#pragma unroll
    for (int l = 0; l < L; ++l) pc.ta[l] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) pc.tgt[c] = 0.0;

    auto bounds = [&](int i, double& lo, double& hi) {
        lo = 0.0;
        hi = 1.0;
        if (i < L) {
            if (afixed) { lo = pc.ta[i]; hi = pc.ta[i]; }
            if (i == 0 && opaque) {
                lo = 1.0;
                if (afixed) hi = 1.0;
            }
        } else if (smooth && i < L + 3) {
            lo = bg[i - L];
            hi = bg[i - L];
        }
    };

    for (;;) {
        // The image descriptor and layer table are uniform: read with scalar
        // loads.  An opaque zero offset keeps those loads inside the loop
        // (served by the scalar cache) instead of letting them be hoisted into
        // ~200 loop-long SGPRs that would spill.
        int zero;
        asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
        // the descriptors do not change during the launch: read through the
        // constant address space at a wave-uniform index, they are scalar loads
        // (the table's address is also read from the argument segment at an
        // opaque offset, so that it is not hoisted into a register that spills)
        const int img = __builtin_amdgcn_readfirstlane(q.img) + zero;
        ImageDescC* dd = as_constant((&args.images)[zero]) + img;
        const DevLayer* lay = (const DevLayer*)dd->layers;
        opaque = dd->opaque != 0;
        smooth = REFINE && dd->smooth != 0;
        // the outer iterate, touched once per outer iteration: a per-lane
        // global slot, [workgroup][variable][lane], its address formed here
        // from the argument read at the opaque offset, so that neither the
        // slot's base nor N precomputed 64-bit addresses stay live across
        // the loop
        double* const xo = (&args.lane_state)[zero] + (long long)blockIdx.x * (4 * kMaxLayers * 64) + lane;
        // the solver constants, read where they are used (scalar loads at the
        // opaque offset) rather than held in scalar registers for the launch
        const SolverParams& sp = (&args.sp)[zero];

        // ---- next pixels for the lanes that finished theirs ----
        // (handed out in batches of at least kFetchBatch lanes while others are
        // still solving: the start of a pixel is a long stretch of code that the
        // whole wave executes whenever any lane needs it)
        unsigned long long needy = __ballot(phase == kPhNeedPixel);
        const unsigned long long retry = __ballot(phase == kPhRetry);
        const bool serve = __popcll(needy | retry) >= kFetchBatch ||
                           __ballot(phase == kPhStart || phase == kPhTrial || phase == kPhOuter || phase == kPhDir) == 0;
        if (!serve) needy = 0;
        long long item = -1;
        while (needy) {
            if (q.next >= q.end) {
                if (q.pending || q.exhausted) break;
                int i2;
                long long b2, e2;
                // the launch arguments at the opaque offset too: claim_chunk's
                // reads of them stay here instead of in launch-long registers
                if (!claim_chunk((&args)[zero], i2, b2, e2)) {
                    q.exhausted = true;
                    break;
                }
                if (i2 == q.img) {
                    q.next = b2;
                    q.end = e2;
                } else {
                    q.pending = true;
                    q.pimg = i2;
                    q.pnext = b2;
                    q.pend = e2;
                    break;
                }
            }
            const long long avail = q.end - q.next;
            const long long rank = __popcll(needy & ((1ull << lane) - 1ull));
            const bool mine = ((needy >> lane) & 1ull) && rank < avail;
            if (mine) item = q.next + rank;
            const unsigned long long took = __ballot(mine);
            q.next += __popcll(took);
            needy &= ~took;
        }
        if (item >= 0 || (serve && phase == kPhRetry)) {
            // a new pixel starts from the warm-start prediction when there is
            // one; a retried pixel from the default start
            if (phase != kPhRetry) pix = dd->list ? dd->list[item] : item;
            warm_try = !REFINE && dd->warm != nullptr && phase != kPhRetry;
            double x0[N], C0[NC], tg0[3], ta0[L];
            load_pixel<L, REFINE>(*dd, lay, pix, tg0, ta0, bg, x0, warm_try);
#pragma unroll
            for (int c = 0; c < 3; ++c) pc.tgt[c] = tg0[c];
#pragma unroll
            for (int l = 0; l < L; ++l) pc.ta[l] = ta0[l];
#pragma unroll
            for (int j = 0; j < NC; ++j) lam[j] = 0.0;
            rho = sp.rho0;
            ocount = 0;
            constraints<L, REFINE, GRAY>(x0, lay, tg0, ta0, C0);
            cn_outer = norm<NC>(C0);
            count = 0;
            head = 0;
            gamma = 1.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                xo[i * 64] = x0[i];
                xc[i] = x0[i];
            }
            phase = kPhStart;
        } else if (phase == kPhNeedPixel && (q.pending || q.exhausted)) {
            phase = kPhWait;
        }

        // ---- a drained wave moves to its pending chunk, or ends ----
        if (__ballot(phase == kPhStart || phase == kPhTrial || phase == kPhOuter || phase == kPhRetry ||
                     phase == kPhDir) == 0) {
            if (!q.pending) {
                q.img = -1;
                return;
            }
            q.img = q.pimg;
            q.next = q.pnext;
            q.end = q.pend;
            q.pending = false;
            phase = kPhNeedPixel;
            continue;
        }

        // ---- the trial point of every lane, projected on the bounds: the
        // start point (a new pixel, an outer restart) or the current point
        // moved t along the direction (a new direction, a backtrack).  One
        // pass for all lanes, rather than one in each branch that sets a
        // trial point; lanes without a pixel compute a value nobody reads ----
        {
            const bool from_start = phase == kPhStart;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                double lo, hi;
                bounds(i, lo, hi);
                S[i] = clampd(from_start ? xc[i] : xc[i] + t * dv[i], lo, hi);
            }
        }

        // ---- one evaluation per lane at the trial point S ----
        ++lc.trips;
        double Ct[PS::NCC];
        const double ft = objective<L, REFINE, GRAY, SPARSE>(S, lay, pc.tgt, pc.ta, lam, rho, Y, Ct,
                                                             SPARSE ? sp.sparse_sigma : 0.0);

        bool direction = false, inner_end = false;
        if (phase == kPhStart) {
            nev = 1;
            f = ft;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                xc[i] = S[i];
                gv[i] = Y[i];
            }
#pragma unroll
            for (int j = 0; j < PS::NCC; ++j) cc[j] = Ct[j];
            direction = true;
        } else if (phase == kPhTrial) {
            ++nev;
            double dphi = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) dphi += gv[i] * (S[i] - xc[i]);
            if (__builtin_isfinite(ft) && ft <= f + 1e-4 * dphi) {
                double sy = 0.0, yy = 0.0;
                bool xstop = true;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    const double xt = S[i], gt = Y[i];
                    const double sv = xt - xc[i];
                    const double yv = gt - gv[i];
                    sy += sv * yv;
                    yy += yv * yv;
                    xstop = xstop && (__builtin_fabs(sv) < sp.xtol * (__builtin_fabs(xt) + __builtin_fabs(xc[i])) * 0.5 ||
                                      xt == xc[i]);
                    xc[i] = xt;
                    gv[i] = gt;
                    S[i] = sv;
                    Y[i] = yv;
                }
                const double fold = f;
                f = ft;
#pragma unroll
                for (int j = 0; j < PS::NCC; ++j) cc[j] = Ct[j];
                if (sy > 2.220446049250313e-16 * yy) {
                    rh = 1.0 / sy;
                    gamma = sy / yy;
                    if constexpr (HIST) {
                        // the pair goes to the ring: appended, or over the
                        // oldest record once the ring is full
                        const int cap = (&args.hist_cap)[zero];
                        int slot;
                        if (count < cap) {
                            slot = head + count;
                            if (slot >= cap) slot -= cap;
                            ++count;
                        } else {
                            slot = head;
                            head = (head + 1 == cap) ? 0 : head + 1;
                        }
                        constexpr int R = kHistRecord(L);
                        double* const rec = (&args.hist)[zero] +
                                            ((long long)blockIdx.x * 64 + lane) * (long long)cap * R +
                                            (long long)slot * R;
#pragma unroll
                        for (int i = 0; i < N; ++i) {
                            rec[i] = S[i];
                            rec[N + i] = Y[i];
                        }
                        rec[2 * N] = rh;
                    } else {
                        count = 1;
                    }
                } else {
                    count = 0;
                    head = 0;
                    gamma = 1.0;
                }
                if (nev >= sp.max_evals || relstop(fold, f, sp.ftol) || xstop) inner_end = true;
                else direction = true;
            } else if (nev >= sp.max_evals || ls + 1 >= sp.max_ls) {
                inner_end = true;
            } else {
                double tn;
                if (dphi < 0.0) {
                    const double denom = 2.0 * (ft - f - dphi);
                    tn = (-dphi * t) / denom;
                } else {
                    tn = 0.5 * t;
                }
                const double lo = 0.1 * t, hi = 0.5 * t;
                tn = (tn >= lo) ? tn : lo;
                tn = (tn > hi) ? hi : tn;
                t = tn;
                ++ls;
            }
        }

        if (direction) {
            // free set and projected gradient; q is then turned into d in place
            freev = 0;
            bool finite = __builtin_isfinite(f);
            double q[N];
            double pgmax = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const double gi = gv[i];
                finite = finite && __builtin_isfinite(gi);
                double lo, hi;
                bounds(i, lo, hi);
                const bool fr = lo < hi && !(xc[i] <= lo && gi > 0.0) && !(xc[i] >= hi && gi < 0.0);
                freev |= fr ? (1u << i) : 0u;
                q[i] = fr ? gi : 0.0;
                const double m = __builtin_fabs(q[i]);
                pgmax = (m > pgmax) ? m : pgmax;
            }
            if (!finite || pgmax == 0.0) {
                inner_end = true;
            } else if constexpr (HIST) {
                ++lc.iters;
                lc.pairs += (unsigned)count;
#pragma unroll
                for (int i = 0; i < N; ++i) dv[i] = q[i];
                dstep = 0;
                phase = kPhDir;
            } else {
                ++lc.iters;
                lc.pairs += (unsigned)count;
                double a = 0.0;
                if (count > 0) {
                    a = rh * tree_dot<N>(S, q);
#pragma unroll
                    for (int i = 0; i < N; ++i) q[i] = q[i] - a * Y[i];
                }
#pragma unroll
                for (int i = 0; i < N; ++i) q[i] = gamma * q[i];
                if (count > 0) {
                    const double b = rh * tree_dot<N>(Y, q);
#pragma unroll
                    for (int i = 0; i < N; ++i) q[i] = q[i] + (a - b) * S[i];
                }
                double gd = 0.0;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    q[i] = ((freev >> i) & 1u) ? -q[i] : 0.0;
                    gd += gv[i] * q[i];
                }
                if (!(gd < 0.0)) {
                    count = 0;
                    gamma = 1.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) q[i] = ((freev >> i) & 1u) ? -gv[i] : 0.0;
                }
                t = 1.0;
                if (count == 0) {
                    double dd = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) dd += q[i] * q[i];
                    const double inv = 1.0 / __builtin_sqrt(dd);
                    t = (inv < 1.0) ? inv : 1.0;
                }
                ls = 0;
#pragma unroll
                for (int i = 0; i < N; ++i) dv[i] = q[i];
                phase = kPhTrial;
            }
        }
        if constexpr (HIST) {
            // a direction with an empty history is q itself (gamma is 1 whenever
            // the history is empty, and 1 q = q exactly): it needs no service
            const unsigned long long pend = __ballot(phase == kPhDir && count > 0);
            unsigned long long done = __ballot(phase == kPhDir && count == 0);
            if (pend) {
                constexpr int R = kHistRecord(L);
                const int cap = (&args.hist_cap)[zero];
                const double* hblock = (&args.hist)[zero] + (long long)blockIdx.x * 64 * (long long)cap * R;
                done |= direction_service<L>(pend, smem + N * 64, hblock, cap, count, head, gamma, dstep);
            }
            if ((done >> lane) & 1ull) {
                double qd[N];
                double gd = 0.0;
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    qd[i] = ((freev >> i) & 1u) ? -dv[i] : 0.0;
                    gd += gv[i] * qd[i];
                }
                if (!(gd < 0.0)) {
                    count = 0;
                    head = 0;
                    gamma = 1.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) qd[i] = ((freev >> i) & 1u) ? -gv[i] : 0.0;
                }
                t = 1.0;
                if (count == 0) {
                    double dd = 0.0;
#pragma unroll
                    for (int i = 0; i < N; ++i) dd += qd[i] * qd[i];
                    const double inv = 1.0 / __builtin_sqrt(dd);
                    t = (inv < 1.0) ? inv : 1.0;
                }
                ls = 0;
#pragma unroll
                for (int i = 0; i < N; ++i) dv[i] = qd[i];
                phase = kPhTrial;
            }
        }

        if (inner_end) phase = kPhOuter;
        const unsigned long long outer_lanes = __ballot(phase == kPhOuter);
        const bool run_outer = __popcll(outer_lanes) >= kOuterBatch ||
                               __ballot(phase == kPhStart || phase == kPhTrial || phase == kPhDir) == 0;
        if (run_outer && phase == kPhOuter) {
            lc.evals += (unsigned)nev;
            lc.outer += 1;
            // the constraint vector at the current point: the composite part
            // kept from its evaluation, the refinement's alpha part alpha_i - t_i
            double Ccur[NC];
#pragma unroll
            for (int j = 0; j < PS::NCC; ++j) Ccur[j] = cc[j];
            if constexpr (REFINE) {
#pragma unroll
                for (int i = 0; i < L; ++i) Ccur[3 + i] = xc[i] - pc.ta[i];
            }
            if constexpr (GRAY) gray_rows(xc + L + 3, Ccur + Shape<L, REFINE, GRAY>::NB);
#pragma unroll
            for (int j = 0; j < NC; ++j) lam[j] = lam[j] - rho * Ccur[j];
            const double ngn = norm<NC>(Ccur);
            if (ngn > sp.gamma * cn_outer) rho = rho * sp.beta;
            // the previous outer iterate: every load issued before the first
            // use, so that the lane waits for one memory round trip, not N
            double xprev[N];
#pragma unroll
            for (int i = 0; i < N; ++i) xprev[i] = xo[i * 64];
            double dx2 = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const double dxv = xc[i] - xprev[i];
                dx2 += dxv * dxv;
            }
#pragma unroll
            for (int i = 0; i < N; ++i) xo[i * 64] = xc[i];
            const bool unchanged = __builtin_sqrt(dx2) < sp.eps;
            const bool satisfied = ngn < sp.eps;
            cn_outer = ngn;
            const bool converged = unchanged && satisfied;
            if (!REFINE && warm_try && !converged && ocount > sp.warm_max_count) {
                // a warm-started attempt that has not met the plain solve's
                // tolerances (|C| < eps and step < eps) by its outer cap: the
                // pixel is solved again from the default start
                if (dd->fallback_count) atomicAdd(dd->fallback_count, 1ull);
                phase = kPhRetry;
            } else if (converged || ocount > sp.max_count) {
#pragma unroll
                for (int l = 0; l < L; ++l) {
                    double* px = dd->out + 4 * ((long long)l * dd->n + pix);
                    px[0] = xc[L + 3 * l];
                    px[1] = xc[L + 3 * l + 1];
                    px[2] = xc[L + 3 * l + 2];
                    px[3] = xc[l];
                    if (dd->alpha_planes) dd->alpha_planes[(long long)l * dd->n + pix] = xc[l];
                }
                ++lc.pixels;
                phase = kPhNeedPixel;
            } else {
                ++ocount;
                count = 0;
                head = 0;
                gamma = 1.0;
                phase = kPhStart;
            }
        }
    }
}

template <int L, bool REFINE, bool HIST, int VAR>
__global__ void __launch_bounds__(64) solve_batch_kernel(BatchArgs args)
{
    __shared__ double smem[persistent_lds_doubles<L, REFINE, VAR>() * 64];
    LaneCounters lc{0, 0, 0, 0, 0, 0};
    WaveQueue q{};
    q.img = -1;
    int i0;
    long long b0, e0;
    if (claim_chunk(args, i0, b0, e0)) {
        q.img = i0;
        q.next = b0;
        q.end = e0;
    }
    // every image of the launch has L layers (the host groups them so)
    if (q.img >= 0) run_images<L, REFINE, HIST, VAR>(args, smem, q, lc);
    unsigned long long evals = lc.evals, iters = lc.iters, outer = lc.outer, pixels = lc.pixels, pairs = lc.pairs;
    for (int off = 32; off > 0; off >>= 1) {
        evals += __shfl_xor(evals, off);
        iters += __shfl_xor(iters, off);
        outer += __shfl_xor(outer, off);
        pixels += __shfl_xor(pixels, off);
        pairs += __shfl_xor(pairs, off);
    }
    if (threadIdx.x == 0) {
        atomicAdd(&args.counters[kEvals], evals);
        atomicAdd(&args.counters[kIterations], iters);
        atomicAdd(&args.counters[kOuter], outer);
        atomicAdd(&args.counters[kPixels], pixels);
        atomicAdd(&args.counters[kPairs], pairs);
        atomicAdd(&args.counters[kSlots], 64ull * lc.trips);
    }
}

template <int L, bool REFINE, bool HIST, int VAR = 0>
hipError_t launch_batch(const BatchArgs& args, hipStream_t stream)
{
    auto kernel = solve_batch_kernel<L, REFINE, HIST, VAR>;
    int device = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0);
    if (e != hipSuccess) return e;
    long long blocks = (long long)cus * (per_cu > 0 ? per_cu : 1);
    // no more workgroups than chunks (each wave takes at least one)
    if (blocks > args.chunk_count) blocks = args.chunk_count;
    if (args.max_blocks > 0 && blocks > args.max_blocks) blocks = args.max_blocks;
    if (blocks > kMaxPersistentBlocks) blocks = kMaxPersistentBlocks;
    // a history-keeping launch has ring storage for hist_blocks blocks
    if (HIST && blocks > args.hist_blocks) blocks = args.hist_blocks;
    if (blocks < 1) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(64), 0, stream, args);
    return hipGetLastError();
}

template <int L>
hipError_t launch_L(bool refine, const PixelArgs& args, hipStream_t stream)
{
    const unsigned blocks = (unsigned)((args.n + 63) / 64);
    if (refine)
        hipLaunchKernelGGL((solve_simple_kernel<L, true, kLbfgsMemory>), dim3(blocks), dim3(64), 0, stream, args);
    else
        hipLaunchKernelGGL((solve_simple_kernel<L, false, kLbfgsMemory>), dim3(blocks), dim3(64), 0, stream, args);
    return hipGetLastError();
}

}  // namespace

}  // namespace ubhip
