// ub_mlp.hip -- the warm-start network: 3x3 RGB patch (27 inputs) -> three
// hidden layers of 128 (GELU) -> 4L sigmoid outputs (alpha, R, G, B per layer,
// the packed first-pass order), trained on the solver's own first-pass
// results.  BF16 operands on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16)
// with FP32 accumulation; FP32 master weights and Adam moments.
//
// Every product is C = A . Bt^T with both operands row-major and contiguous in
// the reduction index k, so each lane fetches its eight BF16 values of an
// operand with one 16-byte load.  Operand lane layout of the instruction:
//   A (32 x 16): lane l holds A[l % 32][8 (l / 32) + j], j = 0..7
//   B (16 x 32): lane l holds B[8 (l / 32) + j][l % 32]
//   C (32 x 32): element r of lane l is C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// Activations and gradients are written by the epilogues in both orientations
// ([sample][feature] for the next layer, [feature][sample] for the weight
// gradient), which keeps every product in that form.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ub_mlp.hpp"

namespace ubhip {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// The logistic function 1 / (1 + e^-x) from the hardware base-2 exponential
// and reciprocal (about 1 ulp each; the products round to BF16 anyway):
// straight-line code, saturating to 0 and 1 for large |x|.
constexpr float kLog2e = 1.4426950408889634f;

__device__ __forceinline__ float logistic(float x)
{
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-kLog2e * x));
}

// GELU, tanh form: 0.5 z (1 + tanh u) with u = c0 (z + c1 z^3), which is
// z * logistic(2u) since 1 + tanh u = 2 logistic(2u).
__device__ __forceinline__ float gelu(float z)
{
    const float u = kGeluC0 * (z + kGeluC1 * z * z * z);
    return z * logistic(2.0f * u);
}

// d/dz of the same: 0.5 (1 + t) + 0.5 z (1 - t^2) u'(z) with t = tanh u,
// 0.5 (1 + t) = s = logistic(2u) and 1 - t^2 = 4 s (1 - s).
__device__ __forceinline__ float gelu_grad(float z)
{
    const float u = kGeluC0 * (z + kGeluC1 * z * z * z);
    const float s = logistic(2.0f * u);
    return s + 2.0f * z * s * (1.0f - s) * kGeluC0 * (1.0f + 3.0f * kGeluC1 * z * z);
}

__device__ __forceinline__ float sigmoid(float z) { return logistic(z); }

// Row loader for a dense BF16 operand.
struct DenseRows {
    const __bf16* p;
    long long ld;
    __device__ __forceinline__ bf16x8 load(long long row, int k) const
    {
        return *reinterpret_cast<const bf16x8*>(p + row * ld + k);
    }
};

// Row loader that gathers the 3x3 clamp-edge RGB patch of a pixel (inputs
// 0..26 = neighbour (dy, dx) in row-major order, channel fastest; 27..31 = 0).
// Training pixels come from a stack of equally sized images.
struct PatchRows {
    const double* rgb;     // [image][h][w][3]
    const long long* idx;  // sample -> global pixel, or nullptr for pixel = base + row
    long long rows;        // valid rows (clamped beyond)
    int w, h;
    long long base;
    __device__ __forceinline__ bf16x8 load(long long row, int k0) const
    {
        if (row >= rows) row = rows - 1;
        const long long g = idx ? idx[row] : base + row;
        const long long n = (long long)w * h;
        const long long img = g / n, p = g - img * n;
        const int y = (int)(p / w), x = (int)(p - (long long)y * w);
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + j;
            float f = 0.0f;
            if (k < 27) {
                const int nb = k / 3, c = k - 3 * nb;
                int yy = y + nb / 3 - 1, xx = x + nb % 3 - 1;
                yy = yy < 0 ? 0 : (yy >= h ? h - 1 : yy);
                xx = xx < 0 ? 0 : (xx >= w ? w - 1 : xx);
                f = (float)(rgb[((img * h + yy) * w + xx) * 3 + c] - 0.5);
            }
            v[j] = (__bf16)f;
        }
        return v;
    }
};

// C[m][n] = sum_k A[m][k] Bt[n][k]; one wave per 32 x 32 block, four waves a
// workgroup (64 x 64).  M and N are multiples of 32, K of 16.
//
// An epilogue returns a per-element value that is summed over the block and
// added to *epi.sum() with one atomic per wave when Epi::kReduce is set.
//
// With gridDim.z > 1 the reduction is split: workgroup z covers k in
// [z K, (z + 1) K) and its epilogue writes a partial product (EpiPartial).
template <class ALoad, class Epi>
__global__ void __launch_bounds__(256) gemm_kernel(ALoad a, DenseRows bt, long long M, int N, int K, Epi epi)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long m0 = (long long)blockIdx.y * 64 + (wave >> 1) * 32;
    const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
    if (m0 >= M || n0 >= N) return;
    const int half = lane >> 5, r32 = lane & 31;
    f32x16 acc = {};
    const int kbeg = blockIdx.z * K;
    for (int k0 = kbeg; k0 < kbeg + K; k0 += 16) {
        const bf16x8 av = a.load(m0 + r32, k0 + 8 * half);
        const bf16x8 bv = bt.load(n0 + r32, k0 + 8 * half);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    float part = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) part += epi(m0 + 8 * (r >> 2) + 4 * half + (r & 3), n0 + r32, acc[r]);
    if constexpr (Epi::kReduce) {
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
        if (lane == 0) atomicAdd(epi.sum(), part);
    }
}

// hidden layer forward: z = acc + b; keep z (FP32) for the backward pass and
// h = GELU(z) in both BF16 orientations.
struct EpiHidden {
    const float* bias;
    float* z;       // [rows][N]
    __bf16* h;      // [rows][N]
    __bf16* ht;     // [N][rows] or nullptr
    long long rows;
    int N;
    static constexpr bool kReduce = false;
    __device__ float* sum() const { return nullptr; }
    __device__ __forceinline__ float operator()(long long m, int n, float v) const
    {
        const float zz = v + bias[n];
        if (z) z[m * N + n] = zz;
        const __bf16 hv = (__bf16)gelu(zz);
        h[m * N + n] = hv;
        if (ht) ht[(long long)n * rows + m] = hv;
        return 0.0f;
    }
};

// Output layer during training: y = sigmoid(z); loss = mean over the valid
// outputs of (y - t)^2; dz = 2 (y - t) y (1 - y) / (rows * outputs).
struct EpiOutputTrain {
    const float* bias;
    const double* targets;  // packed first passes [image][L][n][4]
    const long long* idx;
    long long rows;
    int outputs, layers, w, h;
    float* dz;              // [rows][32] FP32 (bias gradient)
    __bf16* dzb;            // [rows][32]
    __bf16* dzt;            // [32][rows]
    float* loss;
    static constexpr bool kReduce = true;
    __device__ float* sum() const { return loss; }
    __device__ __forceinline__ float operator()(long long m, int n, float v) const
    {
        float d = 0.0f, se = 0.0f;
        if (n < outputs) {
            const float y = sigmoid(v + bias[n]);
            const long long g = idx[m];
            const long long np = (long long)w * h;
            const long long img = g / np, p = g - img * np;
            const int l = n >> 2, c = n & 3;
            const float t = (float)targets[((img * layers + l) * np + p) * 4 + c];
            const float e = y - t;
            d = 2.0f * e * y * (1.0f - y) / (float)(rows * outputs);
            se = e * e / (float)(rows * outputs);
        }
        dz[m * kOutPad + n] = d;
        dzb[m * kOutPad + n] = (__bf16)d;
        dzt[(long long)n * rows + m] = (__bf16)d;
        return se;
    }
};

// Output layer at inference: the prediction, written as a packed first pass.
struct EpiOutputPredict {
    const float* bias;
    double* out;  // [L][n][4]
    long long n, base;
    int outputs;
    static constexpr bool kReduce = false;
    __device__ float* sum() const { return nullptr; }
    __device__ __forceinline__ float operator()(long long m, int j, float v) const
    {
        const long long p = base + m;
        if (j < outputs && m < n - base) out[((long long)(j >> 2) * n + p) * 4 + (j & 3)] = (double)sigmoid(v + bias[j]);
        return 0.0f;
    }
};

// Backward through a hidden layer: dz = (dZ_next W^T) * GELU'(z).
struct EpiBackHidden {
    const float* z;   // [rows][N]
    float* dz;        // [rows][N] FP32 (bias gradient)
    __bf16* dzb;      // [rows][N]
    __bf16* dzt;      // [N][rows]
    long long rows;
    int N;
    static constexpr bool kReduce = false;
    __device__ float* sum() const { return nullptr; }
    __device__ __forceinline__ float operator()(long long m, int n, float v) const
    {
        const float d = v * gelu_grad(z[m * N + n]);
        dz[m * N + n] = d;
        dzb[m * N + n] = (__bf16)d;
        dzt[(long long)n * rows + m] = (__bf16)d;
        return 0.0f;
    }
};

// One split of a split reduction: partial[z][m][n].
struct EpiPartial {
    float* partial;
    long long M;
    int N;
    static constexpr bool kReduce = false;
    __device__ float* sum() const { return nullptr; }
    __device__ __forceinline__ float operator()(long long m, int n, float v) const
    {
        partial[((long long)blockIdx.z * M + m) * N + n] = v;
        return 0.0f;
    }
};

// One gradient reduction of a training step: the split sum of a
// weight-gradient product (out[m][n] = sum over z, in order, of
// partial[z][m][n]) or the column sums of a layer's dz [rows][N] (its bias
// gradient).
struct GradReduce {
    const float* src;  // partial [splits][rows][N], or dz [rows][N]
    float* out;
    long long rows;
    int N;
    int splits;  // > 0: a split sum; 0: column sums
};
constexpr int kGradReduces = 8;  // the four weight and four bias gradients
struct GradReduces {
    GradReduce r[kGradReduces];
    // Adam (Kingma & Ba) on each parameter as its gradient is formed, when
    // update is set: the reductions' outputs cover every parameter once
    Model md;
    const float* grad;  // the gradient vector the outputs lie in
    float lr, b1, b2, eps, c1, c2;
    int update;
};

// Adam on parameter i with gradient g: the FP32 master weight, then both
// BF16 copies.
__device__ __forceinline__ void adam_param(const Model& md, int i, float g, float lr, float b1, float b2, float eps,
                                           float c1, float c2)
{
    const float m = b1 * md.m[i] + (1.0f - b1) * g;
    const float v = b2 * md.v[i] + (1.0f - b2) * g * g;
    md.m[i] = m;
    md.v[i] = v;
    const float w = md.w[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    md.w[i] = w;
    write_bf16(md, i, w);
}

// Every gradient reduction of a step in one launch, each gradient followed
// by its parameter's Adam update: blockIdx.y picks the reduction; a split sum
// takes 256 elements a block, column sums one column a block (256 threads,
// partial sums per wave, then the four waves in order).
__global__ void __launch_bounds__(256) grad_reduce_kernel(GradReduces rs)
{
    const GradReduce g = rs.r[blockIdx.y];
    const int base = int(g.out - rs.grad);  // the parameter index of out[0]
    if (g.splits > 0) {
        const long long mn = g.rows * g.N;
        const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= mn) return;
        float s = 0.0f;
        for (int z = 0; z < g.splits; ++z) s += g.src[(long long)z * mn + i];
        g.out[i] = s;
        if (rs.update) adam_param(rs.md, base + int(i), s, rs.lr, rs.b1, rs.b2, rs.eps, rs.c1, rs.c2);
        return;
    }
    const int n = blockIdx.x;
    if (n >= g.N) return;  // the whole block
    float s = 0.0f;
    for (long long m = threadIdx.x; m < g.rows; m += blockDim.x) s += g.src[m * g.N + n];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float t = part[0] + part[1] + part[2] + part[3];
        g.out[n] = t;
        if (rs.update) adam_param(rs.md, base + n, t, rs.lr, rs.b1, rs.b2, rs.eps, rs.c1, rs.c2);
    }
}

// Input patches of a batch in the [feature][sample] orientation (for dW1).
__global__ void patch_transpose_kernel(PatchRows pr, long long rows, __bf16* xt)
{
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= rows) return;
#pragma unroll
    for (int k0 = 0; k0 < kInPad; k0 += 8) {
        const bf16x8 v = pr.load(m, k0);
#pragma unroll
        for (int j = 0; j < 8; ++j) xt[(long long)(k0 + j) * rows + m] = v[j];
    }
}

// Batch sampling: a SplitMix64 stream per (seed, step, sample).
__device__ __forceinline__ unsigned long long splitmix(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void sample_kernel(long long* idx, int batch, long long pool, unsigned long long seed, int step)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    const unsigned long long r = splitmix(seed ^ splitmix(((unsigned long long)step << 32) | (unsigned)b));
    idx[b] = (long long)(r % (unsigned long long)pool);
}

__global__ void refresh_bf16_kernel(Model md)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParamsPadded) write_bf16(md, i, md.w[i]);
}

__global__ void zero_f32_kernel(float* p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0f;
}

// Inference, all four products of a tile of kPredictRows pixels in one
// workgroup: wave v owns tile rows 32 v .. 32 v + 31 and computes every output
// column of its rows, keeping its hidden activations in LDS (BF16, row-major,
// overwritten layer by layer), so waves share no data and nothing but the
// image and the prediction crosses HBM.  The reduction order of every product
// (k ascending, one 32x32x16 MFMA per step) and the epilogue expressions are
// those of gemm_kernel with EpiHidden and EpiOutputPredict: the prediction is
// the layer-by-layer one, bit for bit.
constexpr int kPredictRows = 128;
constexpr int kActStride = kHidden + 8;  // BF16 per LDS row; the pad spreads rows over banks

__global__ void __launch_bounds__(256) predict_kernel(Model md, const double* rgb, int w, int h, long long n,
                                                      int outputs, double* out)
{
    __shared__ __bf16 act[kPredictRows * kActStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int half = lane >> 5, r32 = lane & 31;
    const long long base = (long long)blockIdx.x * kPredictRows;
    const long long rows = (n - base < kPredictRows) ? n - base : kPredictRows;
    const int m0 = wave * 32;
    __bf16* wa = act + m0 * kActStride;  // this wave's rows

    // first layer: the 3x3 patches (rows past the image repeat its last pixel)
    const PatchRows patches{rgb, nullptr, rows, w, h, base};
    f32x16 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = f32x16{};
#pragma unroll
    for (int k0 = 0; k0 < kInPad; k0 += 16) {
        const bf16x8 av = patches.load(m0 + r32, k0 + 8 * half);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(md.wt[0] + (b * 32 + r32) * kInPad + k0 + 8 * half);
            acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[b], 0, 0, 0);
        }
    }
    for (int l = 0;; ++l) {
        // epilogue of hidden layer l: GELU(z + b) into this wave's rows
        const float* bias = md.w + kOffB[l];
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * half + (r & 3), col = b * 32 + r32;
                wa[row * kActStride + col] = (__bf16)gelu(acc[b][r] + bias[col]);
            }
        __syncthreads();
        if (l == 2) break;
        // hidden layer l + 1
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b] = f32x16{};
        for (int k0 = 0; k0 < kHidden; k0 += 16) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(wa + r32 * kActStride + k0 + 8 * half);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bf16x8 bv =
                    *reinterpret_cast<const bf16x8*>(md.wt[l + 1] + (b * 32 + r32) * kHidden + k0 + 8 * half);
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[b], 0, 0, 0);
            }
        }
        __syncthreads();  // every read of layer l's activations precedes their overwrite
    }
    // output layer (32 columns) and sigmoid, written as a packed first pass
    f32x16 y = {};
    for (int k0 = 0; k0 < kHidden; k0 += 16) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(wa + r32 * kActStride + k0 + 8 * half);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(md.wt[3] + r32 * kHidden + k0 + 8 * half);
        y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, y, 0, 0, 0);
    }
    const float* bias = md.w + kOffB[3];
    const int j = r32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const long long m = m0 + 8 * (r >> 2) + 4 * half + (r & 3);
        const long long p = base + m;
        if (j < outputs && m < rows) out[((long long)(j >> 2) * n + p) * 4 + (j & 3)] = (double)sigmoid(y[r] + bias[j]);
    }
}

template <class ALoad, class Epi>
hipError_t gemm(const ALoad& a, const __bf16* bt, long long M, int N, int K, const Epi& epi, hipStream_t s)
{
    const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
    hipLaunchKernelGGL((gemm_kernel<ALoad, Epi>), grid, dim3(256), 0, s, a, DenseRows{bt, K}, M, N, K, epi);
    return hipGetLastError();
}

// Weight gradient g[M][N] = A . Bt^T with the batch (K) split over
// kGradSplits workgroups along z: the partial products go to `partial`, and
// the returned reduction (run with the step's others by grad_reduce_kernel)
// sums them in a fixed order, so the result does not depend on scheduling.
// A has M rows, Bt N rows, both K long.
GradReduce gemm_weight_grad(const __bf16* a, const __bf16* bt, int M, int N, int K, float* partial, float* g,
                            hipStream_t s)
{
    int splits = kGradSplits;
    while (splits > 1 && (K % (16 * splits)) != 0) splits >>= 1;
    const int ks = K / splits;
    const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64), (unsigned)splits);
    hipLaunchKernelGGL((gemm_kernel<DenseRows, EpiPartial>), grid, dim3(256), 0, s, DenseRows{a, K}, DenseRows{bt, K},
                       (long long)M, N, ks, EpiPartial{partial, M, N});
    return GradReduce{partial, g, (long long)M, N, splits};
}

}  // namespace mlp

using namespace mlp;

hipError_t mlp_refresh(const Model& md, hipStream_t s)
{
    hipLaunchKernelGGL(refresh_bf16_kernel, dim3((kParamsPadded + 255) / 256), dim3(256), 0, s, md);
    return hipGetLastError();
}

hipError_t mlp_step(const Model& md, const TrainBatch& tb, const TrainBuffers& buf, int step, bool update,
                    float lr, float* loss, hipStream_t s)
{
    const long long B = tb.batch;
    hipError_t e;
    hipLaunchKernelGGL(sample_kernel, dim3((tb.batch + 255) / 256), dim3(256), 0, s, buf.idx, tb.batch, tb.pool,
                       tb.seed, step);
    if (tb.fixed_idx) {
        e = hipMemcpyAsync(buf.idx, tb.fixed_idx, sizeof(long long) * B, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, s, loss, 1);
    const PatchRows pr{tb.images, buf.idx, B, tb.w, tb.h, 0};
    // forward
    if ((e = gemm(pr, md.wt[0], B, kHidden, kInPad,
                  EpiHidden{md.w + kOffB[0], buf.z[0], buf.h[0], buf.ht[0], B, kHidden}, s)) != hipSuccess) return e;
    for (int l = 1; l < 3; ++l)
        if ((e = gemm(DenseRows{buf.h[l - 1], kHidden}, md.wt[l], B, kHidden, kHidden,
                      EpiHidden{md.w + kOffB[l], buf.z[l], buf.h[l], buf.ht[l], B, kHidden}, s)) != hipSuccess)
            return e;
    if ((e = gemm(DenseRows{buf.h[2], kHidden}, md.wt[3], B, kOutPad, kHidden,
                  EpiOutputTrain{md.w + kOffB[3], tb.targets, buf.idx, B, tb.outputs, tb.layers, tb.w, tb.h,
                                 buf.dz[3], buf.dzb[3], buf.dzt[3], loss}, s)) != hipSuccess) return e;
    // backward: weight-gradient products (each layer's partials in a slot of
    // their own), the gradient flowing into the layer below, then every
    // weight and bias gradient reduction in one launch
    const size_t slot = size_t(kGradSplits) * kHidden * kHidden;
    GradReduces red{};
    red.r[3] = gemm_weight_grad(buf.ht[2], buf.dzt[3], kHidden, kOutPad, (int)B, buf.partial + 3 * slot,
                                buf.grad + kOffW[3], s);
    for (int l = 2; l >= 0; --l) {
        const int nxt = l + 1;
        const int kin = (nxt == 3) ? kOutPad : kHidden;
        if ((e = gemm(DenseRows{buf.dzb[nxt], kin}, md.wb[nxt], B, kHidden, kin,
                      EpiBackHidden{buf.z[l], buf.dz[l], buf.dzb[l], buf.dzt[l], B, kHidden}, s)) != hipSuccess)
            return e;
        if (l > 0) {
            red.r[l] = gemm_weight_grad(buf.ht[l - 1], buf.dzt[l], kHidden, kHidden, (int)B, buf.partial + l * slot,
                                        buf.grad + kOffW[l], s);
        } else {
            hipLaunchKernelGGL(patch_transpose_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, pr, B,
                               buf.xt);
            red.r[0] = gemm_weight_grad(buf.xt, buf.dzt[0], kInPad, kHidden, (int)B, buf.partial, buf.grad + kOffW[0],
                                        s);
        }
    }
    for (int l = 0; l < 4; ++l) {
        const int N = (l == 3) ? kOutPad : kHidden;
        red.r[4 + l] = GradReduce{buf.dz[l], buf.grad + kOffB[l], B, N, 0};
    }
    // blocks along x: the largest split sum (128 x 128 / 256) or column count
    unsigned bx = 1;
    for (int k = 0; k < kGradReduces; ++k) {
        const GradReduce& g = red.r[k];
        const long long need = g.splits > 0 ? (g.rows * g.N + 255) / 256 : g.N;
        bx = std::max<unsigned>(bx, (unsigned)need);
    }
    red.md = md;
    red.grad = buf.grad;
    red.update = update ? 1 : 0;
    red.b1 = 0.9f;
    red.b2 = 0.999f;
    red.eps = 1e-8f;
    red.lr = lr;
    red.c1 = 1.0f - powf(red.b1, (float)(step + 1));
    red.c2 = 1.0f - powf(red.b2, (float)(step + 1));
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(bx, kGradReduces), dim3(256), 0, s, red);
    return hipGetLastError();
}

hipError_t mlp_predict(const Model& md, const double* rgb, int w, int h, int outputs, double* out, hipStream_t s)
{
    const long long n = (long long)w * h;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(predict_kernel, dim3((unsigned)((n + kPredictRows - 1) / kPredictRows)), dim3(256), 0, s, md,
                       rgb, w, h, n, outputs, out);
    return hipGetLastError();
}

}  // namespace ubhip
