// ub_filter.hip -- colour guided filter (He et al.) on gfx950 with exact box sums.
//
// Semantics: unblending/src/image_processing.cpp:214-273 with the clamp-edge
// box of image_processing.cpp:120-151 / image_processing.hpp:215-222, every
// box mean taken over values quantised to a grid 2^-k:
//   Q(v) = rint(v 2^k),  mean = (sum Q) 2^-k (1 / (2r + 1)^2).
// The integers Q are summed in FP64, where sums stay exact while they are
// within 2^53: k per quantity is chosen from a bound |v| <= 2^e so that the
// longest partial sum the kernels form -- a prefix over P(r) columns of
// (2r + 1)-row window sums, P(r) = 256 for r <= 96 (one-wave strips of at most
// 256 columns) and 1024 beyond (workgroup strips of at most 1024 columns) --
// stays within 2^53: k = 53 - e - ceil(log2(P(r) (2r + 1))).  Bounds:
// guidance statistics (I, I I) from max|I|; a channel's (p, I p) from max|I|
// and max|p|; the coefficients from |a_c| <= max|p| / sqrt(eps) and
// |b| <= max|p| + s3 max|I| max|p| / sqrt(eps) (twice the bound
// |a|_2 <= sd(p) / (2 sqrt(eps)) that holds when the statistics are
// covariances of one data set; coefficients are clamped to the bound, so the
// sums stay exact whatever the data).  Exact sums do not depend on the order
// of summation: the tiling below, and any launch shape the autotuner picks,
// gives the host oracle's result bit for bit.  The per-pixel arithmetic after
// the sums is the reference's, in its order.  No sum is ever -0.0 (the
// accumulators start at +0.0 and only add), as in the oracle's integer sums.
//
// Kernels:
//   bounds   max|I| and max|p| of every channel (one pass over the inputs)
//   pass A   box statistics (9 guidance + 4 per channel) and the per-pixel
//            coefficients a = (Sigma + eps I)^-1 cov_Ip, b = mean_p - a . mean_I
//            (r <= 96: 1 to 3 channels a wave, the guidance statistics
//            recomputed by each wave: reading stored guidance terms at every
//            output pixel of every channel costs more than the recomputation)
//   pass B   box of (a_r, a_g, a_b, b) and q = box(b) + sum_c box(a_c) I_c
// r <= 96: one-wave strips.  A wave owns 64 K consecutive (clamped) columns,
// K per lane, for one channel, and walks RH rows: vertical window sums are
// running sums in registers (the entering and the leaving rows are fetched a
// row ahead); a row's horizontal window sums are differences of the strip's
// prefix sums (a lane's own K columns, then a DPP scan of the lane totals),
// the left end fetched from the lane 2r + 1 columns back by ds_bpermute.  No
// LDS and no barriers.  The waves of one tile (its strip and rows, every
// channel group) run back to back on one XCD, so that they share its L2 for
// the guidance, and each XCD takes a contiguous run of tiles (neighbouring
// strips).
// r > 96: workgroup strips of E threads, one column each, the horizontal
// prefix combined across the workgroup's waves through LDS.
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

// the grid exponent k of a quantity bounded by 2^e for radius r; orc_qscale
__device__ __forceinline__ int qscale(int e, int r)
{
    const long long n = (long long)(r <= kWaveStripMaxRadius ? 256 : 1024) * (2 * r + 1);
    int c = 0;
    while ((1ll << c) < n) ++c;
    const int k = 53 - e - c;
    return k > 1000 ? 1000 : k;
}

// a quantiser: grid 2^-k, values bounded by 2^e
struct Quant {
    double scale;  // 2^k
    double lim;    // 2^(e + k)
    double back;   // 2^-k
    // values known to lie within the bound
    __device__ __forceinline__ double q(double v) const { return __builtin_rint(v * scale); }
    // values clamped to the bound (which bounded values never reach)
    __device__ __forceinline__ double qc(double v) const
    {
        const double t = __builtin_rint(v * scale);
        return __builtin_fmin(__builtin_fmax(t, -lim), lim);
    }
    __device__ __forceinline__ double mean(double s, double weight) const { return (s * back) * weight; }
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

constexpr double kSqrt3Up = 1.7320508075688774;  // sqrt(3) rounded up; the oracle's ORC_SQRT3_UP

// grids of the guidance statistics (I and I I) and of one channel's (p and I p)
__device__ __forceinline__ void stat_quants(const FilterArgs& a, int ch, Quant& qg, Quant& qc)
{
    const int eI = qexp(bound_of(a, 0));
    const int eP = qexp(bound_of(a, 1 + ch));
    qg = make_quant(eI > 2 * eI ? eI : 2 * eI, a.r);
    qc = make_quant(eP > eI + eP ? eP : eI + eP, a.r);
}

// grids of one channel's coefficients: |a_c| <= pmax / sqrt(eps), |b| <= pmax + s3 imax pmax / sqrt(eps)
__device__ __forceinline__ void coeff_quants(const FilterArgs& a, int ch, Quant& qa, Quant& qb)
{
    const double imax = bound_of(a, 0), pmax = bound_of(a, 1 + ch);
    const double ba = pmax * a.isr;
    const double bb = pmax + (kSqrt3Up * imax) * ba;
    qa = make_quant(qexp(ba), a.r);
    qb = make_quant(qexp(bb), a.r);
}

// The per-pixel guidance terms of image_processing.cpp:247-258 from the
// window sums G[0..2] I, G[3..8] I I (00 01 02 11 12 22): the entries of
// (Sigma + eps I)^-1 as cofactor * (1 / det) products (00 01 02 11 12 22),
// then mean_I (3).
__device__ __forceinline__ void guidance_terms(const double* G, const Quant& qg, double weight, double eps,
                                               double gi[9])
{
    const double mI0 = qg.mean(G[0], weight), mI1 = qg.mean(G[1], weight), mI2 = qg.mean(G[2], weight);
    const double m00 = (qg.mean(G[3], weight) - mI0 * mI0) + eps;
    const double m01 = qg.mean(G[4], weight) - mI0 * mI1;
    const double m02 = qg.mean(G[5], weight) - mI0 * mI2;
    const double m11 = (qg.mean(G[6], weight) - mI1 * mI1) + eps;
    const double m12 = qg.mean(G[7], weight) - mI1 * mI2;
    const double m22 = (qg.mean(G[8], weight) - mI2 * mI2) + eps;
    const double c00 = m11 * m22 - m12 * m12;
    const double c01 = m02 * m12 - m01 * m22;
    const double c02 = m01 * m12 - m02 * m11;
    const double c11 = m00 * m22 - m02 * m02;
    const double c12 = m01 * m02 - m00 * m12;
    const double c22 = m00 * m11 - m01 * m01;
    const double det = m00 * c00 + m01 * c01 + m02 * c02;
    const double id = 1.0 / det;
    gi[0] = c00 * id;
    gi[1] = c01 * id;
    gi[2] = c02 * id;
    gi[3] = c11 * id;
    gi[4] = c12 * id;
    gi[5] = c22 * id;
    gi[6] = mI0;
    gi[7] = mI1;
    gi[8] = mI2;
}

// a = (Sigma + eps I)^-1 cov_Ip and b = mean_p - a . mean_I
// (image_processing.cpp:258-265) from the guidance terms and a channel's
// window sums P[0] p, P[1..3] I p
__device__ __forceinline__ void channel_coefficients(const double* gi, const double* P, const Quant& qc,
                                                     double weight, double out[4])
{
    const double mp = qc.mean(P[0], weight);
    const double v0 = qc.mean(P[1], weight) - gi[6] * mp;
    const double v1 = qc.mean(P[2], weight) - gi[7] * mp;
    const double v2 = qc.mean(P[3], weight) - gi[8] * mp;
    const double ar = gi[0] * v0 + gi[1] * v1 + gi[2] * v2;
    const double ag = gi[1] * v0 + gi[3] * v1 + gi[4] * v2;
    const double ab = gi[2] * v0 + gi[4] * v1 + gi[5] * v2;
    out[0] = ar;
    out[1] = ag;
    out[2] = ab;
    out[3] = ((mp - ar * gi[6]) - ag * gi[7]) - ab * gi[8];
}

// the 9 guidance statistics of one pixel, quantised
__device__ __forceinline__ void guidance_stats(const Quant& qg, double I0, double I1, double I2, double* s)
{
    s[0] = qg.q(I0);
    s[1] = qg.q(I1);
    s[2] = qg.q(I2);
    s[3] = qg.q(I0 * I0);
    s[4] = qg.q(I0 * I1);
    s[5] = qg.q(I0 * I2);
    s[6] = qg.q(I1 * I1);
    s[7] = qg.q(I1 * I2);
    s[8] = qg.q(I2 * I2);
}

// a channel's 4 statistics of one pixel, quantised
__device__ __forceinline__ void channel_stats(const Quant& qc, double I0, double I1, double I2, double p, double* s)
{
    s[0] = qc.q(p);
    s[1] = qc.q(I0 * p);
    s[2] = qc.q(I1 * p);
    s[3] = qc.q(I2 * p);
}

// ---- bounds: max|I| (slot 0) and max|p_c| (slot 1 + c) ----
// One source after another (the guidance as 3n values, then each channel),
// contiguous sources read as 16-byte pairs, four in flight per thread.
__device__ __forceinline__ double absmax2(double m, double2 v)
{
    return __builtin_fmax(m, __builtin_fmax(__builtin_fabs(v.x), __builtin_fabs(v.y)));
}

__global__ void __launch_bounds__(256) bounds_kernel(BoundsArgs a)
{
    const long long n = (long long)a.w * a.h;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nth = (long long)gridDim.x * blockDim.x;
    for (int src = -1; src < a.channels; ++src) {
        const double* base = src < 0 ? a.rgb : a.in[src].base;
        const long long count = src < 0 ? 3 * n : n;
        const int stride = src < 0 ? 1 : a.in[src].stride;
        double m = 0.0;
        if (stride == 1 && ((unsigned long long)base & 15) == 0) {
            const double2* p2 = reinterpret_cast<const double2*>(base);
            const long long pairs = count >> 1;
            long long i = tid;
            for (; i + 3 * nth < pairs; i += 4 * nth) {
                const double2 v0 = p2[i], v1 = p2[i + nth], v2 = p2[i + 2 * nth], v3 = p2[i + 3 * nth];
                m = absmax2(absmax2(m, v0), v1);
                m = absmax2(absmax2(m, v2), v3);
            }
            for (; i < pairs; i += nth) m = absmax2(m, p2[i]);
            if ((count & 1) && tid == 0) m = __builtin_fmax(m, __builtin_fabs(base[count - 1]));
        } else {
            for (long long i = tid; i < count; i += nth) m = __builtin_fmax(m, __builtin_fabs(base[i * stride]));
        }
        for (int off = 32; off > 0; off >>= 1) m = __builtin_fmax(m, __shfl_xor(m, off));
        // one atomic per workgroup and source; non-negative doubles order as their bit patterns
        __shared__ double part[4];
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            m = __builtin_fmax(__builtin_fmax(part[0], part[1]), __builtin_fmax(part[2], part[3]));
            atomicMax(&a.bounds[src + 1], (unsigned long long)__builtin_bit_cast(long long, m));
        }
        __syncthreads();
    }
}

// ---- r = 0: the identity ----
__global__ void copy_kernel(FilterArgs a)
{
    const long long n = (long long)a.w * a.h;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int g = 0; g < a.channels; ++g) a.q[g][i] = a.in[g].base[i * a.in[g].stride];
}

// v of the lane the DPP control selects (+0.0 where it selects none)
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dppd(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// inclusive prefix sum across the 64 lanes of a wave: shifts by 1, 2, 4, 8
// within rows of 16 lanes, then the row broadcasts of lanes 15 and 31 (GFX9 DPP)
__device__ __forceinline__ double wave_scan(double v)
{
    v += dppd<0x111, 0xf>(v);  // row_shr:1
    v += dppd<0x112, 0xf>(v);  // row_shr:2
    v += dppd<0x114, 0xf>(v);  // row_shr:4
    v += dppd<0x118, 0xf>(v);  // row_shr:8
    v += dppd<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dppd<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// v of lane `byte_index / 4` (any lane of the wave)
__device__ __forceinline__ double lane_fetch(double v, int byte_index)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(byte_index, (int)(unsigned)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_ds_bpermute(byte_index, (int)(b >> 32));
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// ===========================================================================
// One-wave strips (r <= 96)
// ===========================================================================

// Two consecutive doubles at an 8-byte-aligned address: one 16-byte access.
typedef double dpair __attribute__((ext_vector_type(2), aligned(8)));

// K values of a channel from the columns cols[] (clamped) of a row
template <int K>
__device__ __forceinline__ void load_channel(const Channel& ch, long long rowbase, const int* cols, double* v, int step)
{
#pragma unroll
    for (int j = 0; j < K; ++j) v[j * step] = ch.base[(rowbase + cols[j]) * ch.stride];
}

template <int K>
__device__ __forceinline__ void load_rgb(const double* rgb, long long rowbase, const int* cols, double* v, int step)
{
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const double* p = rgb + 3 * (rowbase + cols[j]);
        v[j * step] = p[0];
        v[j * step + 1] = p[1];
        v[j * step + 2] = p[2];
    }
}

// Pass A, the guidance wave of a tile: raw inputs (I_r, I_g, I_b), the 9
// guidance statistics.
struct GuideSrc {
    static constexpr int NR = 3;
    static constexpr int NS = 9;
    const double* rgb;
    int w;
    double eps, weight;
    Quant qg;
    __device__ __forceinline__ void init(const FilterArgs& a)
    {
        rgb = a.rgb;
        w = a.w;
        eps = a.eps;
        weight = 1.0 / (double)((2 * a.r + 1) * (2 * a.r + 1));
        Quant qc;
        stat_quants(a, 0, qg, qc);
    }
    template <int K>
    __device__ __forceinline__ void load(long long rowbase, const int* cols, double v[][NR]) const
    {
        load_rgb<K>(rgb, rowbase, cols, &v[0][0], NR);
    }
    __device__ __forceinline__ void stats(const double* v, double* s) const { guidance_stats(qg, v[0], v[1], v[2], s); }
};

// Pass A, a channel wave of a tile: raw inputs (I_r, I_g, I_b, p), the
// statistics p, I_r p, I_g p, I_b p; its coefficients combine them with the
// guidance terms of the tile's guidance wave.
struct ChanSrc {
    static constexpr int NR = 4;
    static constexpr int NS = 4;
    const double* rgb;
    Channel in;
    double* ab;  // this channel's 4 planes
    long long n;
    int w;
    double weight;
    Quant qc;
    __device__ __forceinline__ void init(const FilterArgs& a, int ch)
    {
        rgb = a.rgb;
        in = a.in[ch];
        n = (long long)a.w * a.h;
        ab = a.ab + (long long)(4 * ch) * n;
        w = a.w;
        weight = 1.0 / (double)((2 * a.r + 1) * (2 * a.r + 1));
        Quant qg;
        stat_quants(a, ch, qg, qc);
    }
    template <int K>
    __device__ __forceinline__ void load(long long rowbase, const int* cols, double v[][NR]) const
    {
        load_rgb<K>(rgb, rowbase, cols, &v[0][0], NR);
        load_channel<K>(in, rowbase, cols, &v[0][3], NR);
    }
    __device__ __forceinline__ void stats(const double* v, double* s) const
    {
        channel_stats(qc, v[0], v[1], v[2], v[3], s);
    }
    __device__ __forceinline__ void out(long long pix, const double* gi, const double* H) const
    {
        double c[4];
        channel_coefficients(gi, H, qc, weight, c);
#pragma unroll
        for (int k = 0; k < 4; ++k) ab[(long long)k * n + pix] = c[k];
    }
    // the coefficients of K consecutive output pixels pix0 ..., 16-byte stores
    template <int K>
    __device__ __forceinline__ void out_run(long long pix0, const double c[][4]) const
    {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double* p = ab + (long long)k * n + pix0;
#pragma unroll
            for (int j = 0; j < K; j += 2) *(dpair*)(p + j) = dpair{c[j][k], c[j + 1][k]};
        }
    }
};

// Pass B of one channel: raw inputs and statistics (a_r, a_g, a_b, b),
// output q = box(b) + box(a_r) I_r + box(a_g) I_g + box(a_b) I_b
// (image_processing.cpp:267-270, that order).
struct WaveB {
    static constexpr int NR = 4;
    static constexpr int NS = 4;
    const double* rgb;
    const double* ab;
    double* q;
    long long n;
    int w;
    double weight;
    Quant qa, qb;
    __device__ __forceinline__ void init(const FilterArgs& a, int ch)
    {
        rgb = a.rgb;
        n = (long long)a.w * a.h;
        ab = a.ab + (long long)(4 * ch) * n;
        q = a.q[ch];
        w = a.w;
        weight = 1.0 / (double)((2 * a.r + 1) * (2 * a.r + 1));
        coeff_quants(a, ch, qa, qb);
    }
    template <int K>
    __device__ __forceinline__ void load(long long rowbase, const int* cols, double v[][NR]) const
    {
        // in-plane offsets fit 32 bits (n <= (2^31 - 1) / 16 on the device)
        const int row = (int)rowbase;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double* pk = ab + (long long)k * n;
#pragma unroll
            for (int j = 0; j < K; ++j) v[j][k] = pk[row + cols[j]];
        }
    }
    __device__ __forceinline__ void stats(const double* v, double* s) const
    {
        s[0] = qa.qc(v[0]);
        s[1] = qa.qc(v[1]);
        s[2] = qa.qc(v[2]);
        s[3] = qb.qc(v[3]);
    }
    __device__ __forceinline__ void out(long long pix, const double* H) const
    {
        const double I0 = rgb[3 * pix], I1 = rgb[3 * pix + 1], I2 = rgb[3 * pix + 2];
        double v = qb.mean(H[3], weight);
        v = v + qa.mean(H[0], weight) * I0;
        v = v + qa.mean(H[1], weight) * I1;
        v = v + qa.mean(H[2], weight) * I2;
        q[pix] = v;
    }
    // K consecutive output pixels pix0 ...: 16-byte guidance loads and stores
    template <int K>
    __device__ __forceinline__ void out_run(long long pix0, const double H[][NS]) const
    {
        double I[3 * K];
        const double* g = rgb + 3 * pix0;
#pragma unroll
        for (int i = 0; i < 3 * K; i += 2) {
            const dpair d = *(const dpair*)(g + i);
            I[i] = d.x;
            I[i + 1] = d.y;
        }
        double v[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            double t = qb.mean(H[j][3], weight);
            t = t + qa.mean(H[j][0], weight) * I[3 * j];
            t = t + qa.mean(H[j][1], weight) * I[3 * j + 1];
            t = t + qa.mean(H[j][2], weight) * I[3 * j + 2];
            v[j] = t;
        }
#pragma unroll
        for (int j = 0; j < K; j += 2) *(dpair*)(q + pix0 + j) = dpair{v[j], v[j + 1]};
    }
};

// One wave's share of a strip: input columns X0 - r + [0, 64 K), K per lane;
// output columns X0 + [0, TW), TW = 64 K - 2r.  Holds the running vertical
// window sums of the wave's statistics and the next row's inputs.  RHO =
// (2r + 1) mod K fixes, per slot j of a lane, which slot of which lane holds
// the prefix just left of its window.
template <class Src, int K, int RHO>
struct StripRows {
    static constexpr int NS = Src::NS;
    static constexpr int NR = Src::NR;
    double V[K][NS];
    double vin[K][NR], vout[K][NR];
    int cols[K];  // the lane's input columns, clamped to the image
    int lane, r, w, h, base;

    __device__ __forceinline__ void setup(int X0, int r_, int w_, int h_)
    {
        lane = threadIdx.x & 63;
        r = r_;
        w = w_;
        h = h_;
#pragma unroll
        for (int j = 0; j < K; ++j) cols[j] = clampi(X0 - r + lane * K + j, 0, w - 1);
        // the prefix just left of slot j's window, column lane K + j - (2r + 1):
        // lane lane - base - (j < RHO), slot (j - RHO) mod K
        base = (2 * r + 1 - RHO) / K;
    }
    // window sums of row y0 (rows y0 - r .. y0 + r)
    __device__ __forceinline__ void warmup(const Src& src, int y0)
    {
#pragma unroll
        for (int j = 0; j < K; ++j)
#pragma unroll
            for (int k = 0; k < NS; ++k) V[j][k] = 0.0;
        for (int dy = -r; dy <= r; ++dy) {
            double v[K][NR];
            src.template load<K>((long long)clampi(y0 + dy, 0, h - 1) * w, cols, v);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double s[NS];
                src.stats(v[j], s);
#pragma unroll
                for (int k = 0; k < NS; ++k) V[j][k] += s[k];
            }
        }
    }
    // to row y (from y - 1 unless first), then fetch the inputs of row y + 1's update
    __device__ __forceinline__ void advance(const Src& src, int y, bool first, bool more)
    {
        if (!first) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double si[NS], so[NS];
                src.stats(vin[j], si);
                src.stats(vout[j], so);
#pragma unroll
                for (int k = 0; k < NS; ++k) V[j][k] = V[j][k] + (si[k] - so[k]);
            }
        }
        if (more) {
            src.template load<K>((long long)clampi(y + 1 + r, 0, h - 1) * w, cols, vin);
            src.template load<K>((long long)clampi(y - r, 0, h - 1) * w, cols, vout);
        }
    }
    // window sums of the row: H = S(c) - S(c - 2r - 1), S the strip's prefix
    __device__ __forceinline__ void window(double H[K][NS]) const
    {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            double S[K];
            S[0] = V[0][k];
#pragma unroll
            for (int j = 1; j < K; ++j) S[j] = S[j - 1] + V[j][k];
            const double T = S[K - 1];
            const double ex = wave_scan(T) - T;
#pragma unroll
            for (int j = 0; j < K; ++j) S[j] = ex + S[j];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int from = lane - base - (j < RHO ? 1 : 0);
                const double left = lane_fetch(S[(j - RHO + K) % K], from << 2);
                H[j][k] = S[j] - (from >= 0 ? left : 0.0);
            }
        }
    }
};

// Tile of a launch: block b runs on XCD b mod 8; the channel groups of a tile
// are consecutive blocks of one XCD, and each XCD takes a contiguous run of
// tiles (neighbouring strips of the same rows).  Returns false past the end.
struct TilePos {
    int tile, group, strip, chunk;
};
__device__ __forceinline__ bool tile_of_block(int tiles, int strips, int groups, TilePos& t)
{
    const int xcd = blockIdx.x & 7;
    const int local = blockIdx.x >> 3;
    t.group = local % groups;
    const int tl = local / groups;
    const int per_xcd = (tiles + 7) >> 3;
    t.tile = xcd * per_xcd + tl;
    if (tl >= per_xcd || t.tile >= tiles) return false;
    t.strip = t.tile % strips;
    t.chunk = t.tile / strips;
    return true;
}

// Pass A, r <= 96: a workgroup of 1 + S waves per tile and channel group.
// Wave 0 keeps the guidance statistics and, each row, writes the guidance
// terms of its output columns to LDS; waves 1..S each keep one channel's
// statistics and, after the row's barrier, combine them with the guidance
// terms into the channel's coefficients.  The channel waves load the same
// guidance rows as wave 0, from the CU's cache.
template <int K, int RHO>
__global__ void __launch_bounds__(512) coeff_tile_kernel(FilterArgs a, int TW, int RH, int strips, int tiles, int S)
{
    __shared__ double gt[2][9][K][64];
    const int groups = (a.channels + S - 1) / S;
    TilePos t;
    if (!tile_of_block(tiles, strips, groups, t)) return;  // the whole workgroup
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int X0 = t.strip * TW;
    const int y0 = t.chunk * RH;
    const int y1 = min(a.h, y0 + RH);
    const int r = a.r;

    if (wave == 0) {
        GuideSrc src;
        src.init(a);
        StripRows<GuideSrc, K, RHO> st;
        st.setup(X0, r, a.w, a.h);
        st.warmup(src, y0);
        int buf = 0;
        for (int y = y0; y < y1; ++y) {
            st.advance(src, y, y == y0, y + 1 < y1);
            double H[K][9];
            st.window(H);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double gi[9];
                guidance_terms(H[j], src.qg, src.weight, src.eps, gi);
#pragma unroll
                for (int k = 0; k < 9; ++k) gt[buf][k][j][lane] = gi[k];
            }
            __syncthreads();
            buf ^= 1;
        }
        return;
    }
    const int ch = t.group * S + wave - 1;
    if (ch >= a.channels) {
        for (int y = y0; y < y1; ++y) __syncthreads();  // an idle wave keeps the barrier count
        return;
    }
    ChanSrc src;
    src.init(a, ch);
    StripRows<ChanSrc, K, RHO> st;
    st.setup(X0, r, a.w, a.h);
    st.warmup(src, y0);
    int buf = 0;
    for (int y = y0; y < y1; ++y) {
        st.advance(src, y, y == y0, y + 1 < y1);
        double H[K][4];
        st.window(H);
        __syncthreads();
        const int c0 = lane * K;  // the lane's first strip column; column c's window is strip columns c - 2r .. c
        const int x0 = X0 + c0 - 2 * r;
        if (c0 >= 2 * r && c0 + K - 1 - 2 * r < TW && x0 + K - 1 < a.w) {
            // all K columns are output columns
            double cf[K][4];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                double gi[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) gi[k] = gt[buf][k][j][lane];
                channel_coefficients(gi, H[j], src.qc, src.weight, cf[j]);
            }
            src.template out_run<K>((long long)y * a.w + x0, cf);
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int c = c0 + j;
                const int x = x0 + j;
                if (c >= 2 * r && c - 2 * r < TW && x < a.w) {
                    double gi[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) gi[k] = gt[buf][k][j][lane];
                    src.out((long long)y * a.w + x, gi, H[j]);
                }
            }
        }
        buf ^= 1;
    }
}

// Pass B, r <= 96: a workgroup of S waves per tile and channel group, one
// channel each (the waves read the same guidance at their output pixels,
// from the CU's cache).
template <int K, int RHO>
__global__ void __launch_bounds__(512) output_tile_kernel(FilterArgs a, int TW, int RH, int strips, int tiles, int S)
{
    const int groups = (a.channels + S - 1) / S;
    TilePos t;
    if (!tile_of_block(tiles, strips, groups, t)) return;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int ch = t.group * S + wave;
    if (ch >= a.channels) return;
    const int X0 = t.strip * TW;
    const int y0 = t.chunk * RH;
    const int y1 = min(a.h, y0 + RH);
    const int r = a.r;
    WaveB src;
    src.init(a, ch);
    StripRows<WaveB, K, RHO> st;
    st.setup(X0, r, a.w, a.h);
    st.warmup(src, y0);
    for (int y = y0; y < y1; ++y) {
        st.advance(src, y, y == y0, y + 1 < y1);
        double H[K][4];
        st.window(H);
        const int c0 = lane * K;
        const int x0 = X0 + c0 - 2 * r;
        if (c0 >= 2 * r && c0 + K - 1 - 2 * r < TW && x0 + K - 1 < a.w) {
            src.template out_run<K>((long long)y * a.w + x0, H);
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int c = c0 + j;
                const int x = x0 + j;
                if (c >= 2 * r && c - 2 * r < TW && x < a.w) src.out((long long)y * a.w + x, H[j]);
            }
        }
    }
}

// ===========================================================================
// Workgroup strips (r > 96)
// ===========================================================================

// Pass A of a group of G channels (first_channel ...): raw inputs I (3) and
// p (G); statistics I (3), I I (6), then per channel p, I_r p, I_g p, I_b p.
template <int G>
struct PassA {
    static constexpr int NR = 3 + G;
    static constexpr int NS = 9 + 4 * G;
    FilterArgs a;
    Quant qg, qc[G];
    double weight;

    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int g = 0; g < G; ++g) stat_quants(a, a.first_channel + g, qg, qc[g]);
        weight = 1.0 / (double)((2 * a.r + 1) * (2 * a.r + 1));
    }
    __device__ __forceinline__ void raw(long long pix, double* v) const
    {
        v[0] = a.rgb[3 * pix];
        v[1] = a.rgb[3 * pix + 1];
        v[2] = a.rgb[3 * pix + 2];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const Channel& c = a.in[g];  // the launch's channels start at in[0]
            v[3 + g] = c.base[pix * c.stride];
        }
    }
    __device__ __forceinline__ void stats(const double* v, double* s) const
    {
        guidance_stats(qg, v[0], v[1], v[2], s);
#pragma unroll
        for (int g = 0; g < G; ++g) channel_stats(qc[g], v[0], v[1], v[2], v[3 + g], s + 9 + 4 * g);
    }
    __device__ __forceinline__ void out(long long pix, long long n, const double* H) const
    {
        double gi[9];
        guidance_terms(H, qg, weight, a.eps, gi);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            double c[4];
            channel_coefficients(gi, H + 9 + 4 * g, qc[g], weight, c);
            double* ab = a.ab + (long long)(4 * (a.first_channel + g)) * n;
#pragma unroll
            for (int k = 0; k < 4; ++k) ab[(long long)k * n + pix] = c[k];
        }
    }
};

// Pass B of a group: raw inputs and statistics (a_r, a_g, a_b, b) per channel.
template <int G>
struct PassB {
    static constexpr int NR = 4 * G;
    static constexpr int NS = 4 * G;
    FilterArgs a;
    Quant qa[G], qb[G];
    double weight;

    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int g = 0; g < G; ++g) coeff_quants(a, a.first_channel + g, qa[g], qb[g]);
        weight = 1.0 / (double)((2 * a.r + 1) * (2 * a.r + 1));
    }
    __device__ __forceinline__ void raw(long long pix, double* v) const
    {
        const long long n = (long long)a.w * a.h;
        const double* ab = a.ab + (long long)(4 * a.first_channel) * n;
#pragma unroll
        for (int k = 0; k < 4 * G; ++k) v[k] = ab[(long long)k * n + pix];
    }
    __device__ __forceinline__ void stats(const double* v, double* s) const
    {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            s[4 * g + 0] = qa[g].qc(v[4 * g + 0]);
            s[4 * g + 1] = qa[g].qc(v[4 * g + 1]);
            s[4 * g + 2] = qa[g].qc(v[4 * g + 2]);
            s[4 * g + 3] = qb[g].qc(v[4 * g + 3]);
        }
    }
    __device__ __forceinline__ void out(long long pix, long long, const double* H) const
    {
        const double I0 = a.rgb[3 * pix], I1 = a.rgb[3 * pix + 1], I2 = a.rgb[3 * pix + 2];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            double q = qb[g].mean(H[4 * g + 3], weight);
            q = q + qa[g].mean(H[4 * g + 0], weight) * I0;
            q = q + qa[g].mean(H[4 * g + 1], weight) * I1;
            q = q + qa[g].mean(H[4 * g + 2], weight) * I2;
            a.q[g][pix] = q;  // the launch's channels start at q[0]
        }
    }
};

// A workgroup of E threads owns E consecutive (clamped) columns, of which the
// middle TW = E - 2r are its output columns, and walks RH rows: vertical
// window sums in registers, a row's horizontal window sums the differences
// of a workgroup-wide inclusive prefix (wave scans, one LDS exchange).
template <class Src, int E>
__global__ void __launch_bounds__(E) strip_kernel(Src src, int TW, int RH)
{
    constexpr int NS = Src::NS;
    constexpr int NR = Src::NR;
    constexpr int NW = E / 64;
    // statistics exchanged through LDS per round (48 KiB of prefixes)
    constexpr int KC = (6144 / E) < NS ? (6144 / E) : NS;
    __shared__ double pref[KC][E];
    __shared__ double tot[KC][NW];

    src.init();
    const FilterArgs& a = src.a;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int r = a.r;
    const long long n = (long long)a.w * a.h;
    const int col = clampi(blockIdx.x * TW - r + t, 0, a.w - 1);
    const int y0 = blockIdx.y * RH;
    const int y1 = min(a.h, y0 + RH);

    double V[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) V[k] = 0.0;
    for (int dy = -r; dy <= r; ++dy) {
        double v[NR], s[NS];
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
        double H[NS];
#pragma unroll
        for (int c0 = 0; c0 < NS; c0 += KC) {
            double P[KC];
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                if (c0 + j < NS) {
                    P[j] = wave_scan(V[c0 + j]);
                    pref[j][t] = P[j];
                    if (lane == 63) tot[j][wave] = P[j];
                }
            }
            __syncthreads();
            if (writer) {
#pragma unroll
                for (int j = 0; j < KC; ++j) {
                    if (c0 + j < NS) {
                        double own = P[j], lo = 0.0;
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
            double si[NS], so[NS];
            src.stats(vin, si);
            src.stats(vout, so);
#pragma unroll
            for (int k = 0; k < NS; ++k) V[k] = V[k] + (si[k] - so[k]);
        }
    }
}

// ===========================================================================
// Launch geometry
// ===========================================================================

struct WaveGeometry {
    int K, TW, RH, strips, chunks, tiles;
};

// One-wave strips: 64 K columns (K = 4 unless block asks for 128 columns and
// the radius leaves 64 output columns), rows per strip so that the launch has
// about 4096 channel waves, at least 32; rows overrides it (the autotuner's knob).
bool wave_geometry(int w, int h, int r, int channels, int block, int rows, WaveGeometry& g)
{
    if (r < 1 || r > kWaveStripMaxRadius) return false;
    g.K = (block == 128 && 128 - 2 * r >= 64) ? 2 : 4;
    g.TW = 64 * g.K - 2 * r;
    g.strips = (w + g.TW - 1) / g.TW;
    int rh;
    if (rows > 0) {
        rh = rows;
    } else {
        const long long per_row = (long long)g.strips * channels;
        const long long want = (4096 + per_row - 1) / per_row;
        rh = (int)((h + want - 1) / want);
        rh = rh < 32 ? 32 : rh;
    }
    g.RH = rh;
    g.chunks = (h + rh - 1) / rh;
    g.tiles = g.strips * g.chunks;
    return true;
}

constexpr int kTileChannels = 7;  // channel waves of one pass-A workgroup (plus the guidance wave)
constexpr int kOutChannels = 8;   // channel waves of one pass-B workgroup

template <bool PASS_A, int K, int RHO>
hipError_t launch_tiles_k(const FilterArgs& a, const WaveGeometry& g, hipStream_t s)
{
    const int S = PASS_A ? (a.channels < kTileChannels ? a.channels : kTileChannels)
                         : (a.channels < kOutChannels ? a.channels : kOutChannels);
    const int groups = (a.channels + S - 1) / S;
    const long long blocks = 8ll * ((g.tiles + 7) / 8) * groups;
    if (blocks > 0x7fffffffll) return hipErrorInvalidValue;
    if (PASS_A)
        hipLaunchKernelGGL((coeff_tile_kernel<K, RHO>), dim3((unsigned)blocks), dim3(64 * (1 + S)), 0, s, a, g.TW,
                           g.RH, g.strips, g.tiles, S);
    else
        hipLaunchKernelGGL((output_tile_kernel<K, RHO>), dim3((unsigned)blocks), dim3(64 * S), 0, s, a, g.TW, g.RH,
                           g.strips, g.tiles, S);
    return hipGetLastError();
}

template <bool PASS_A>
hipError_t launch_tiles(const FilterArgs& a, hipStream_t s)
{
    WaveGeometry g;
    if (!wave_geometry(a.w, a.h, a.r, a.channels, a.block, a.rows, g)) return hipErrorInvalidValue;
    const int rho = (2 * a.r + 1) % g.K;
    if (g.K == 2) return launch_tiles_k<PASS_A, 2, 1>(a, g, s);
    return rho == 1 ? launch_tiles_k<PASS_A, 4, 1>(a, g, s) : launch_tiles_k<PASS_A, 4, 3>(a, g, s);
}

struct Geometry {
    int E, TW, RH, strips, chunks;
};

// Workgroup strips: width E (its middle TW = E - 2r columns are output
// columns), by default the smallest width whose output part is at least half
// of it, and RH aiming at ~2048 workgroups, at least 2r + 2 rows.
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
hipError_t launch_pass(const FilterArgs& args, hipStream_t s)
{
    if (args.w <= 0 || args.h <= 0) return hipSuccess;
    if (args.channels < 1 || args.channels > kMaxChannels) return hipErrorInvalidValue;
    if (args.r < 0 || args.r > kMaxFilterRadius) return hipErrorInvalidValue;
    if (args.r == 0) {
        if (PASS_A) return hipSuccess;  // the identity needs no coefficients
        const long long n = (long long)args.w * args.h;
        hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, args);
        return hipGetLastError();
    }
    if (args.r <= kWaveStripMaxRadius) return launch_tiles<PASS_A>(args, s);
    Geometry geo;
    if (!geometry(args.w, args.h, args.r, args.block, args.rows, geo)) return hipErrorInvalidValue;
    // 1024-thread strips of pass B one channel a launch: with three their
    // 12 running sums spill (128 registers a lane at that size), while pass A
    // keeps three, since each launch recomputes the 9 guidance sums
    const int gmax = (geo.E == 1024 && !PASS_A) ? 1 : kMaxGroup;
    for (int g0 = 0; g0 < args.channels; g0 += gmax) {
        // the launch's channels are moved to the front of in[] and q[], so that
        // the kernels index those arrays with compile-time indices (a run-time
        // index into the kernel argument would copy it to scratch memory);
        // first_channel still places the group's bounds and coefficient planes
        FilterArgs a = args;
        a.first_channel = g0;
        a.groups = args.channels - g0 < gmax ? args.channels - g0 : gmax;
        for (int c = 0; c < a.groups; ++c) {
            a.in[c] = args.in[g0 + c];
            a.q[c] = args.q[g0 + c];
        }
        hipError_t e;
        switch (geo.E) {
        case 256: e = launch_e<PASS_A, 256>(a, geo, s); break;
        case 512: e = launch_e<PASS_A, 512>(a, geo, s); break;
        default: e = launch_e<PASS_A, 1024>(a, geo, s); break;
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

hipError_t launch_filter_bounds(const BoundsArgs& args, hipStream_t stream)
{
    if (args.w <= 0 || args.h <= 0) return hipSuccess;
    if (args.channels < 1 || args.channels > kMaxChannels) return hipErrorInvalidValue;
    const long long n = (long long)args.w * args.h;
    long long blocks = (3 * n + 2047) / 2048;  // at least 4 pairs of the guidance a thread
    if (blocks > 1024) blocks = 1024;
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
