// ub_filter.hip -- colour guided filter (He et al.) on gfx950, FP64.
//
// Semantics: unblending/src/image_processing.cpp:214-273 with the clamp-edge
// box of image_processing.cpp:120-151 / image_processing.hpp:215-222.
// Two passes, each one strip kernel:
//   pass A  reads the RGB guidance and G input channels, forms the 9 guidance
//           statistics and 4 per-channel statistics (p, I_r p, I_g p, I_b p),
//           box-filters them, and writes a = (Sigma + eps I)^-1 cov_Ip and
//           b = mean_p - a . mean_I  (4 planes a channel).
//   pass B  box-filters (a_r, a_g, a_b, b) and writes
//           q = box(b) + box(a_r) I_r + box(a_g) I_g + box(a_b) I_b.
// A workgroup of E threads owns E consecutive (clamped) image columns, of
// which the middle TW = E - 2r are its output columns, and walks a chunk of
// rows top to bottom.  Each thread keeps the vertical window sums of all
// statistics of its column in registers (add the entering row, subtract the
// leaving row); the horizontal window is a difference of a block-wide
// inclusive prefix sum (wave scan by lane shuffles, wave totals through LDS).
// Clamp-edge handling is exact: out-of-range rows and columns load the edge
// pixel, as the reference's crop() of the sample index does.
#include "ub_kernels.hpp"

namespace ubhip {
namespace {

constexpr int kChunk = 8;  // statistics exchanged through LDS per round

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <int G, bool PASS_A>
struct Stats {
    static constexpr int N = PASS_A ? 9 + 4 * G : 4 * G;
};

template <int G, bool PASS_A>
__device__ __forceinline__ void load_row(const FilterArgs& a, long long n, int y, int x, double* v)
{
    const long long pix = (long long)y * a.w + x;
    if constexpr (PASS_A) {
        const double I0 = a.rgb[3 * pix], I1 = a.rgb[3 * pix + 1], I2 = a.rgb[3 * pix + 2];
        v[0] = I0;
        v[1] = I1;
        v[2] = I2;
        v[3] = I0 * I0;
        v[4] = I0 * I1;
        v[5] = I0 * I2;
        v[6] = I1 * I1;
        v[7] = I1 * I2;
        v[8] = I2 * I2;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const double p = a.in[g].base[pix * a.in[g].stride];
            v[9 + 4 * g] = p;
            v[10 + 4 * g] = I0 * p;
            v[11 + 4 * g] = I1 * p;
            v[12 + 4 * g] = I2 * p;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4 * G; ++k) v[k] = a.ab[(long long)k * n + pix];
    }
}

template <int G, bool PASS_A, int E>
__global__ void __launch_bounds__(E) strip_kernel(FilterArgs a, int TW, int RH)
{
    constexpr int NS = Stats<G, PASS_A>::N;
    constexpr int NW = E / 64;
    __shared__ double pref[kChunk][E];
    __shared__ double tot[kChunk][NW];

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
        double v[NS];
        load_row<G, PASS_A>(a, n, clampi(y0 + dy, 0, a.h - 1), col, v);
#pragma unroll
        for (int k = 0; k < NS; ++k) V[k] += v[k];
    }

    const int ox = blockIdx.x * TW + t - 2 * r;  // output column of this thread
    const bool writer = t >= 2 * r && t - 2 * r < TW && ox < a.w;
    const double inv = 1.0 / (double)((2 * r + 1) * (2 * r + 1));

    for (int y = y0; y < y1; ++y) {
        double H[NS];
#pragma unroll
        for (int c0 = 0; c0 < NS; c0 += kChunk) {
            double s[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
                if (c0 + j < NS) {
                    double v = V[c0 + j];
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const double u = __shfl_up(v, off);
                        if (lane >= off) v += u;
                    }
                    s[j] = v;
                    if (lane == 63) tot[j][wave] = v;
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
                if (c0 + j < NS) {
                    double off = 0.0;
                    for (int w = 0; w < wave; ++w) off += tot[j][w];
                    s[j] += off;
                    pref[j][t] = s[j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
                if (c0 + j < NS) {
                    const int src = t - 2 * r - 1;
                    const double lo = (src >= 0) ? pref[j][src] : 0.0;
                    H[c0 + j] = s[j] - lo;
                }
            }
            __syncthreads();
        }

        if (writer) {
            const long long pix = (long long)y * a.w + ox;
            double m[NS];
#pragma unroll
            for (int k = 0; k < NS; ++k) m[k] = H[k] * inv;
            if constexpr (PASS_A) {
                const double mI0 = m[0], mI1 = m[1], mI2 = m[2];
                const double m00 = (m[3] - mI0 * mI0) + a.eps, m01 = m[4] - mI0 * mI1, m02 = m[5] - mI0 * mI2;
                const double m11 = (m[6] - mI1 * mI1) + a.eps, m12 = m[7] - mI1 * mI2;
                const double m22 = (m[8] - mI2 * mI2) + a.eps;
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
                    const double mp = m[9 + 4 * g];
                    const double v0 = m[10 + 4 * g] - mI0 * mp;
                    const double v1 = m[11 + 4 * g] - mI1 * mp;
                    const double v2 = m[12 + 4 * g] - mI2 * mp;
                    const double ar = (c00 * id) * v0 + (c01 * id) * v1 + (c02 * id) * v2;
                    const double ag = (c01 * id) * v0 + (c11 * id) * v1 + (c12 * id) * v2;
                    const double ab = (c02 * id) * v0 + (c12 * id) * v1 + (c22 * id) * v2;
                    const double b = ((mp - ar * mI0) - ag * mI1) - ab * mI2;
                    a.ab[(long long)(4 * g + 0) * n + pix] = ar;
                    a.ab[(long long)(4 * g + 1) * n + pix] = ag;
                    a.ab[(long long)(4 * g + 2) * n + pix] = ab;
                    a.ab[(long long)(4 * g + 3) * n + pix] = b;
                }
            } else {
                const double I0 = a.rgb[3 * pix], I1 = a.rgb[3 * pix + 1], I2 = a.rgb[3 * pix + 2];
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    double q = m[4 * g + 3];
                    q = q + m[4 * g + 0] * I0;
                    q = q + m[4 * g + 1] * I1;
                    q = q + m[4 * g + 2] * I2;
                    a.q[g][pix] = q;
                }
            }
        }

        if (y + 1 < y1) {
            double vin[NS], vout[NS];
            load_row<G, PASS_A>(a, n, clampi(y + 1 + r, 0, a.h - 1), col, vin);
            load_row<G, PASS_A>(a, n, clampi(y - r, 0, a.h - 1), col, vout);
#pragma unroll
            for (int k = 0; k < NS; ++k) V[k] = (V[k] + vin[k]) - vout[k];
        }
    }
}

struct Geometry {
    int E, TW, RH, strips, chunks;
};

// Strip geometry: workgroup width E (its middle TW = E - 2r columns are
// output columns) and rows per workgroup RH.  By default E is the smallest
// width whose output part is at least half of it and RH aims at ~1024
// workgroups, at least 4r + 2 rows (the vertical warm-up is 2r + 1 rows).
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
        const int want = (1024 + g.strips - 1) / g.strips;
        rh = (h + want - 1) / want;
        rh = rh < 4 * r + 2 ? 4 * r + 2 : rh;
        rh = rh < 32 ? 32 : rh;
    }
    g.RH = rh;
    g.chunks = (h + rh - 1) / rh;
    return true;
}

template <bool PASS_A, int E, int G>
void launch_eg(const FilterArgs& a, const Geometry& geo, hipStream_t s)
{
    hipLaunchKernelGGL((strip_kernel<G, PASS_A, E>), dim3(geo.strips, geo.chunks), dim3(E), 0, s, a, geo.TW, geo.RH);
}

template <bool PASS_A, int E>
hipError_t launch_e(const FilterArgs& a, const Geometry& geo, hipStream_t s)
{
    switch (a.groups) {
    case 1: launch_eg<PASS_A, E, 1>(a, geo, s); break;
    case 2: launch_eg<PASS_A, E, 2>(a, geo, s); break;
    case 3: launch_eg<PASS_A, E, 3>(a, geo, s); break;
    case 4: launch_eg<PASS_A, E, 4>(a, geo, s); break;
    case 5: launch_eg<PASS_A, E, 5>(a, geo, s); break;
    case 6: launch_eg<PASS_A, E, 6>(a, geo, s); break;
    case 7: launch_eg<PASS_A, E, 7>(a, geo, s); break;
    case 8: launch_eg<PASS_A, E, 8>(a, geo, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool PASS_A>
hipError_t launch_pass(const FilterArgs& a, hipStream_t s)
{
    Geometry geo;
    if (a.w <= 0 || a.h <= 0) return hipSuccess;
    if (!geometry(a.w, a.h, a.r, a.block, a.rows, geo)) return hipErrorInvalidValue;
    switch (geo.E) {
    case 256: return launch_e<PASS_A, 256>(a, geo, s);
    case 512: return launch_e<PASS_A, 512>(a, geo, s);
    default: return launch_e<PASS_A, 1024>(a, geo, s);
    }
}

}  // namespace

hipError_t launch_filter_coefficients(const FilterArgs& args, hipStream_t stream)
{
    return launch_pass<true>(args, stream);
}

hipError_t launch_filter_output(const FilterArgs& args, hipStream_t stream)
{
    return launch_pass<false>(args, stream);
}

}  // namespace ubhip
