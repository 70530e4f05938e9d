// ub_filter.hip -- colour guided filter (He et al.) on gfx950 with exact box sums.
//
// Semantics: unblending/src/image_processing.cpp:214-273 with the clamp-edge
// box of image_processing.cpp:120-151 / image_processing.hpp:215-222, every
// box mean taken over values quantised to a grid 2^-k and summed in 64-bit
// integers:
//   Q(v) = rint(v 2^k),  mean = ldexp((double) sum Q, -k) * (1 / (2r + 1)^2),
// with k per quantity from a bound |v| < 2^E such that every window sum stays
// below 2^62 (guidance statistics I, I I from max|I|; a channel's p, I p from
// max|I| and max|p|; the coefficients from the closed-form bounds
// |a| < 7 max|I| max|p| / eps, |b| < max|p| + 21 max|I|^2 max|p| / eps).
// Integer sums are exact, so the statistics do not depend on the order of
// summation: the tiling below (and any launch shape the autotuner picks)
// gives the host oracle's result bit for bit.  The per-pixel arithmetic after
// the sums is the reference's, in its order.
//
// Kernels:
//   bounds   max|I| and max|p| of every channel (one pass over the inputs)
//   pass A   per channel group: box statistics (9 guidance + 4 per channel),
//            a = (Sigma + eps I)^-1 cov_Ip and b = mean_p - a . mean_I
//   pass B   per channel group: box of (a_r, a_g, a_b, b) and
//            q = box(b) + box(a_r) I_r + box(a_g) I_g + box(a_b) I_b
// A strip workgroup of E threads owns E consecutive (clamped) columns, of
// which the middle TW = E - 2r are its output columns, and walks RH rows.
// Each thread keeps its column's vertical window sums in registers (add the
// entering row, subtract the leaving row, whose inputs are fetched one row
// ahead); a row's horizontal window sums are differences of a workgroup-wide
// inclusive prefix sum (wave scans, then one LDS exchange for all statistics).
// r = 0 is the identity (the reference's box returns its input, so q = p).
#include "ub_kernels.hpp"

namespace ubhip {
namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// smallest e with |x| < 2^e (zero and subnormals: -1022); the oracle's orc_qexp
__host__ __device__ __forceinline__ int qexp(double x)
{
    unsigned long long b = (unsigned long long)__builtin_bit_cast(long long, x);
    const int ex = (int)((b >> 52) & 0x7ff);
    return ex == 0 ? -1022 : ex - 1022;
}

// the grid exponent k of a quantity bounded by 2^e in windows of radius r; orc_qscale
__device__ __forceinline__ int qscale(int e, int r)
{
    const long long n = (long long)(2 * r + 1) * (2 * r + 1);
    int c = 0;
    while ((1ll << c) < n) ++c;
    const int k = 62 - e - c;
    return k > 1000 ? 1000 : k;
}

// a quantiser: grid 2^-k, values bounded by 2^e (clamped there, which bounded
// values never reach)
struct Quant {
    double scale;  // 2^k
    double lim;    // 2^(e + k)
    double back;   // 2^-k
    __device__ __forceinline__ long long q(double v) const
    {
        double t = __builtin_rint(v * scale);
        t = t > lim ? lim : (t < -lim ? -lim : t);
        return (long long)t;
    }
    __device__ __forceinline__ double mean(long long s, double weight) const { return ((double)s * back) * weight; }
};

__device__ __forceinline__ Quant make_quant(int e, int r)
{
    const int k = qscale(e, r);
    Quant qt;
    qt.scale = __builtin_ldexp(1.0, k);
    qt.lim = __builtin_ldexp(1.0, e + k);
    qt.back = __builtin_ldexp(1.0, -k);
    return qt;
}

__device__ __forceinline__ double bound_of(const FilterArgs& a, int slot)
{
    return __builtin_bit_cast(double, (long long)a.bounds[slot]);
}

// ---- bounds: max|I| (slot 0) and max|p_c| (slot 1 + c), all channels in one pass ----
__global__ void __launch_bounds__(256) bounds_kernel(BoundsArgs a)
{
    const long long n = (long long)a.w * a.h;
    double mi = 0.0, mp[kMaxChannels];
#pragma unroll
    for (int g = 0; g < kMaxChannels; ++g) mp[g] = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) mi = __builtin_fmax(mi, __builtin_fabs(a.rgb[3 * i + c]));
#pragma unroll
        for (int g = 0; g < kMaxChannels; ++g)
            if (g < a.channels) mp[g] = __builtin_fmax(mp[g], __builtin_fabs(a.in[g].base[i * a.in[g].stride]));
    }
    for (int off = 32; off > 0; off >>= 1) {
        mi = __builtin_fmax(mi, __shfl_xor(mi, off));
#pragma unroll
        for (int g = 0; g < kMaxChannels; ++g) mp[g] = __builtin_fmax(mp[g], __shfl_xor(mp[g], off));
    }
    if ((threadIdx.x & 63) == 0) {
        // non-negative doubles order as their bit patterns
        atomicMax(&a.bounds[0], (unsigned long long)__builtin_bit_cast(long long, mi));
#pragma unroll
        for (int g = 0; g < kMaxChannels; ++g)
            if (g < a.channels) atomicMax(&a.bounds[1 + g], (unsigned long long)__builtin_bit_cast(long long, mp[g]));
    }
}

// ---- r = 0: the identity ----
__global__ void copy_kernel(FilterArgs a)
{
    const long long n = (long long)a.w * a.h;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int g = 0; g < kMaxGroup; ++g)
        if (g < a.groups) a.q[g][i] = a.in[g].base[i * a.in[g].stride];
}

// ---- the sources of the two passes ----

// Pass A: raw inputs I (3) and p (G); statistics I (3), I I (6), then per
// channel p, I_r p, I_g p, I_b p.
template <int G>
struct PassA {
    static constexpr int NR = 3 + G;
    static constexpr int NS = 9 + 4 * G;
    FilterArgs a;
    Quant qg, qc[G];
    double weight;

    __device__ __forceinline__ void init()
    {
        const double imax = bound_of(a, 0);
        const int eI = qexp(imax);
        qg = make_quant(eI > 2 * eI ? eI : 2 * eI, a.r);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int eP = qexp(bound_of(a, 1 + a.first_channel + g));
            qc[g] = make_quant(eP > eI + eP ? eP : eI + eP, a.r);
        }
        weight = 1.0 / (double)((2 * a.r + 1) * (2 * a.r + 1));
    }
    __device__ __forceinline__ void raw(long long pix, double* v) const
    {
        v[0] = a.rgb[3 * pix];
        v[1] = a.rgb[3 * pix + 1];
        v[2] = a.rgb[3 * pix + 2];
#pragma unroll
        for (int g = 0; g < G; ++g) v[3 + g] = a.in[g].base[pix * a.in[g].stride];
    }
    __device__ __forceinline__ void stats(const double* v, long long* s) const
    {
        const double I0 = v[0], I1 = v[1], I2 = v[2];
        s[0] = qg.q(I0);
        s[1] = qg.q(I1);
        s[2] = qg.q(I2);
        s[3] = qg.q(I0 * I0);
        s[4] = qg.q(I0 * I1);
        s[5] = qg.q(I0 * I2);
        s[6] = qg.q(I1 * I1);
        s[7] = qg.q(I1 * I2);
        s[8] = qg.q(I2 * I2);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double p = v[3 + g];
            s[9 + 4 * g] = qc[g].q(p);
            s[10 + 4 * g] = qc[g].q(I0 * p);
            s[11 + 4 * g] = qc[g].q(I1 * p);
            s[12 + 4 * g] = qc[g].q(I2 * p);
        }
    }
    __device__ __forceinline__ void out(long long pix, long long n, const long long* H) const
    {
        const double mI0 = qg.mean(H[0], weight), mI1 = qg.mean(H[1], weight), mI2 = qg.mean(H[2], weight);
        const double m00 = (qg.mean(H[3], weight) - mI0 * mI0) + a.eps;
        const double m01 = qg.mean(H[4], weight) - mI0 * mI1;
        const double m02 = qg.mean(H[5], weight) - mI0 * mI2;
        const double m11 = (qg.mean(H[6], weight) - mI1 * mI1) + a.eps;
        const double m12 = qg.mean(H[7], weight) - mI1 * mI2;
        const double m22 = (qg.mean(H[8], weight) - mI2 * mI2) + a.eps;
        const double c00 = m11 * m22 - m12 * m12;
        const double c01 = m02 * m12 - m01 * m22;
        const double c02 = m01 * m12 - m02 * m11;
        const double c11 = m00 * m22 - m02 * m02;
        const double c12 = m01 * m02 - m00 * m12;
        const double c22 = m00 * m11 - m01 * m01;
        const double det = m00 * c00 + m01 * c01 + m02 * c02;
        const double id = 1.0 / det;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double mp = qc[g].mean(H[9 + 4 * g], weight);
            const double v0 = qc[g].mean(H[10 + 4 * g], weight) - mI0 * mp;
            const double v1 = qc[g].mean(H[11 + 4 * g], weight) - mI1 * mp;
            const double v2 = qc[g].mean(H[12 + 4 * g], weight) - mI2 * mp;
            const double ar = (c00 * id) * v0 + (c01 * id) * v1 + (c02 * id) * v2;
            const double ag = (c01 * id) * v0 + (c11 * id) * v1 + (c12 * id) * v2;
            const double ab = (c02 * id) * v0 + (c12 * id) * v1 + (c22 * id) * v2;
            const double b = ((mp - ar * mI0) - ag * mI1) - ab * mI2;
            a.ab[(long long)(4 * g + 0) * n + pix] = ar;
            a.ab[(long long)(4 * g + 1) * n + pix] = ag;
            a.ab[(long long)(4 * g + 2) * n + pix] = ab;
            a.ab[(long long)(4 * g + 3) * n + pix] = b;
        }
    }
};

// Pass B: raw inputs and statistics (a_r, a_g, a_b, b) per channel.
template <int G>
struct PassB {
    static constexpr int NR = 4 * G;
    static constexpr int NS = 4 * G;
    FilterArgs a;
    Quant qa[G], qb[G];
    double weight;

    __device__ __forceinline__ void init()
    {
        const double imax = bound_of(a, 0);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double pmax = bound_of(a, 1 + a.first_channel + g);
            qa[g] = make_quant(qexp(7.0 * imax * pmax / a.eps), a.r);
            qb[g] = make_quant(qexp(pmax + 21.0 * imax * imax * pmax / a.eps), a.r);
        }
        weight = 1.0 / (double)((2 * a.r + 1) * (2 * a.r + 1));
    }
    __device__ __forceinline__ void raw(long long pix, double* v) const
    {
        const long long n = (long long)a.w * a.h;
#pragma unroll
        for (int k = 0; k < 4 * G; ++k) v[k] = a.ab[(long long)k * n + pix];
    }
    __device__ __forceinline__ void stats(const double* v, long long* s) const
    {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            s[4 * g + 0] = qa[g].q(v[4 * g + 0]);
            s[4 * g + 1] = qa[g].q(v[4 * g + 1]);
            s[4 * g + 2] = qa[g].q(v[4 * g + 2]);
            s[4 * g + 3] = qb[g].q(v[4 * g + 3]);
        }
    }
    __device__ __forceinline__ void out(long long pix, long long, const long long* H) const
    {
        const double I0 = a.rgb[3 * pix], I1 = a.rgb[3 * pix + 1], I2 = a.rgb[3 * pix + 2];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            double q = qb[g].mean(H[4 * g + 3], weight);
            q = q + qa[g].mean(H[4 * g + 0], weight) * I0;
            q = q + qa[g].mean(H[4 * g + 1], weight) * I1;
            q = q + qa[g].mean(H[4 * g + 2], weight) * I2;
            a.q[g][pix] = q;
        }
    }
};

// v of the lane the DPP control selects (0 where it selects none), both halves
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ long long dpp64(long long v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v & 0xffffffffll), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, ROW_MASK, 0xf, false);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// inclusive prefix sum of a 64-bit integer across the 64 lanes of a wave:
// shifts by 1, 2, 4, 8 within rows of 16 lanes, then the row broadcasts of
// lanes 15 and 31 (GFX9 DPP)
__device__ __forceinline__ long long wave_scan(long long v, int)
{
    v += dpp64<0x111, 0xf>(v);  // row_shr:1
    v += dpp64<0x112, 0xf>(v);  // row_shr:2
    v += dpp64<0x114, 0xf>(v);  // row_shr:4
    v += dpp64<0x118, 0xf>(v);  // row_shr:8
    v += dpp64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

template <class Src, int E>
__global__ void __launch_bounds__(E) strip_kernel(Src src, int TW, int RH)
{
    constexpr int NS = Src::NS;
    constexpr int NR = Src::NR;
    constexpr int NW = E / 64;
    // statistics exchanged through LDS per round (48 KiB of prefixes)
    constexpr int KC = (6144 / E) < NS ? (6144 / E) : NS;
    __shared__ long long pref[KC][E];
    __shared__ long long tot[KC][NW];

    src.init();
    const FilterArgs& a = src.a;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int r = a.r;
    const long long n = (long long)a.w * a.h;
    const int col = clampi(blockIdx.x * TW - r + t, 0, a.w - 1);
    const int y0 = blockIdx.y * RH;
    const int y1 = min(a.h, y0 + RH);

    long long V[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) V[k] = 0;
    for (int dy = -r; dy <= r; ++dy) {
        double v[NR];
        long long s[NS];
        src.raw((long long)clampi(y0 + dy, 0, a.h - 1) * a.w + col, v);
        src.stats(v, s);
#pragma unroll
        for (int k = 0; k < NS; ++k) V[k] += s[k];
    }

    const int ox = blockIdx.x * TW + t - 2 * r;  // output column of this thread
    const bool writer = t >= 2 * r && t - 2 * r < TW && ox < a.w;
    const int from = t - 2 * r - 1;              // the window's left neighbour in the strip
    const int fwave = from >= 0 ? from >> 6 : 0;

    for (int y = y0; y < y1; ++y) {
        // inputs of the next row's update, fetched before this row's exchange
        double vin[NR], vout[NR];
        const bool more = y + 1 < y1;
        if (more) {
            src.raw((long long)clampi(y + 1 + r, 0, a.h - 1) * a.w + col, vin);
            src.raw((long long)clampi(y - r, 0, a.h - 1) * a.w + col, vout);
        }

        // horizontal window sums: differences of the strip's prefix sums
        long long H[NS];
#pragma unroll
        for (int c0 = 0; c0 < NS; c0 += KC) {
            long long P[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                if (c0 + j < NS) {
                    P[j] = wave_scan(V[c0 + j], lane);
                    pref[j][t] = P[j];
                    if (lane == 63) tot[j][wave] = P[j];
                }
            }
            __syncthreads();
            if (writer) {
#pragma unroll
                for (int j = 0; j < KC; ++j) {
                    if (c0 + j < NS) {
                        long long own = P[j], lo = 0;
                        for (int w = 0; w < wave; ++w) own += tot[j][w];
                        if (from >= 0) {
                            lo = pref[j][from];
                            for (int w = 0; w < fwave; ++w) lo += tot[j][w];
                        }
                        H[c0 + j] = own - lo;
                    }
                }
            }
            __syncthreads();
        }
        if (writer) src.out((long long)y * a.w + ox, n, H);

        if (more) {
            long long si[NS], so[NS];
            src.stats(vin, si);
            src.stats(vout, so);
#pragma unroll
            for (int k = 0; k < NS; ++k) V[k] = (V[k] + si[k]) - so[k];
        }
    }
}

struct Geometry {
    int E, TW, RH, strips, chunks;
};

// Strip geometry: workgroup width E (its middle TW = E - 2r columns are
// output columns) and rows per workgroup RH.  By default E is the smallest
// width whose output part is at least half of it and RH aims at ~2048
// workgroups, at least 2r + 2 rows (the vertical warm-up is 2r + 1 rows).
// block / rows override them (the autotuner's knobs).
bool geometry(int w, int h, int r, int block, int rows, Geometry& g)
{
    if (r < 0 || r > kMaxFilterRadius) return false;
    if (block == 256 || block == 512 || block == 1024) g.E = block;
    else if (256 - 2 * r >= 128) g.E = 256;
    else if (512 - 2 * r >= 256) g.E = 512;
    else g.E = 1024;
    if (g.E - 2 * r < 1) g.E = 1024;
    g.TW = g.E - 2 * r;
    g.strips = (w + g.TW - 1) / g.TW;
    int rh;
    if (rows > 0) {
        rh = rows;
    } else {
        const int want = (2048 + g.strips - 1) / g.strips;
        rh = (h + want - 1) / want;
        rh = rh < 2 * r + 2 ? 2 * r + 2 : rh;
        rh = rh < 16 ? 16 : rh;
    }
    g.RH = rh;
    g.chunks = (h + rh - 1) / rh;
    return true;
}

template <class Src, int E>
void launch_src(const FilterArgs& a, const Geometry& geo, hipStream_t s)
{
    Src src{};
    src.a = a;
    hipLaunchKernelGGL((strip_kernel<Src, E>), dim3(geo.strips, geo.chunks), dim3(E), 0, s, src, geo.TW, geo.RH);
}

template <bool PASS_A, int E>
hipError_t launch_e(const FilterArgs& a, const Geometry& geo, hipStream_t s)
{
    switch (a.groups) {
    case 1: PASS_A ? launch_src<PassA<1>, E>(a, geo, s) : launch_src<PassB<1>, E>(a, geo, s); break;
    case 2: PASS_A ? launch_src<PassA<2>, E>(a, geo, s) : launch_src<PassB<2>, E>(a, geo, s); break;
    case 3: PASS_A ? launch_src<PassA<3>, E>(a, geo, s) : launch_src<PassB<3>, E>(a, geo, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool PASS_A>
hipError_t launch_pass(const FilterArgs& a, hipStream_t s)
{
    Geometry geo;
    if (a.w <= 0 || a.h <= 0) return hipSuccess;
    if (a.groups < 1 || a.groups > kMaxGroup) return hipErrorInvalidValue;
    if (a.r == 0) {
        if (PASS_A) return hipSuccess;  // the identity needs no coefficients
        const long long n = (long long)a.w * a.h;
        hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    if (!geometry(a.w, a.h, a.r, a.block, a.rows, geo)) return hipErrorInvalidValue;
    switch (geo.E) {
    case 256: return launch_e<PASS_A, 256>(a, geo, s);
    case 512: return launch_e<PASS_A, 512>(a, geo, s);
    default: return launch_e<PASS_A, 1024>(a, geo, s);
    }
}

}  // namespace

hipError_t launch_filter_bounds(const BoundsArgs& args, hipStream_t stream)
{
    if (args.w <= 0 || args.h <= 0) return hipSuccess;
    if (args.channels < 1 || args.channels > kMaxChannels) return hipErrorInvalidValue;
    const long long n = (long long)args.w * args.h;
    long long blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(bounds_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, args);
    return hipGetLastError();
}

hipError_t launch_filter_coefficients(const FilterArgs& args, hipStream_t stream)
{
    return launch_pass<true>(args, stream);
}

hipError_t launch_filter_output(const FilterArgs& args, hipStream_t stream)
{
    return launch_pass<false>(args, stream);
}

}  // namespace ubhip
