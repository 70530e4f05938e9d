// ub_mlp.hpp -- shapes, parameter layout and launch interface of the
// warm-start network (ub_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ubhip {
namespace mlp {

constexpr int kInputs = 27;   // 3x3 neighbourhood x RGB
constexpr int kInPad = 32;
constexpr int kHidden = 128;  // three hidden layers
constexpr int kOutPad = 32;   // outputs 4L <= 32, padded
constexpr float kGeluC0 = 0.7978845608028654f;  // sqrt(2 / pi), tanh form of GELU
constexpr float kGeluC1 = 0.044715f;

// FP32 parameter vector: W1 [32][128], b1, W2 [128][128], b2, W3 [128][128],
// b3, W4 [128][32], b4 -- weights [in][out] row-major.
constexpr int kIn[4] = {kInPad, kHidden, kHidden, kHidden};
constexpr int kOut[4] = {kHidden, kHidden, kHidden, kOutPad};
constexpr int kOffW[4] = {0, 4224, 20736, 37248};
constexpr int kOffB[4] = {4096, 20608, 37120, 41344};
constexpr int kParamsPadded = 41376;
constexpr int kWeightsBf16 = 4096 + 16384 + 16384 + 4096;  // per orientation

// Parameters a network with `outputs` outputs uses (the rest are zero pads).
constexpr int params_used(int outputs)
{
    return kInputs * kHidden + kHidden + 2 * (kHidden * kHidden + kHidden) + kHidden * outputs + outputs;
}

// A model: FP32 master weights and Adam moments, BF16 copies of each weight
// matrix in both orientations (wt = W^T [out][in] for the forward products,
// wb = W [in][out] for the backward ones).
struct Model {
    float* w;
    float* m;
    float* v;
    __bf16* wt[4];
    __bf16* wb[4];
};

inline size_t model_bytes() { return 3 * sizeof(float) * kParamsPadded + 2 * sizeof(uint16_t) * kWeightsBf16; }

inline Model model_view(void* base)
{
    Model md;
    float* f = static_cast<float*>(base);
    md.w = f;
    md.m = f + kParamsPadded;
    md.v = f + 2 * kParamsPadded;
    __bf16* b = reinterpret_cast<__bf16*>(f + 3 * kParamsPadded);
    for (int l = 0; l < 4; ++l) {
        md.wt[l] = b;
        b += kIn[l] * kOut[l];
    }
    for (int l = 0; l < 4; ++l) {
        md.wb[l] = b;
        b += kIn[l] * kOut[l];
    }
    return md;
}

// Writes parameter i's BF16 copies (biases have none).
__device__ __forceinline__ void write_bf16(const Model& md, int i, float w)
{
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int r = i - kOffW[l];
        if (r >= 0 && r < kIn[l] * kOut[l]) {
            const int row = r / kOut[l], col = r - row * kOut[l];
            md.wb[l][r] = (__bf16)w;
            md.wt[l][col * kIn[l] + row] = (__bf16)w;
        }
    }
}

// Training samples: pixels of a stack of equally sized images with their
// solved first passes as targets.
struct TrainBatch {
    const double* images;   // [count][h][w][3]
    const double* targets;  // [count][L][h*w][4]
    int w, h, layers, outputs, batch;
    long long pool;         // count * h * w
    unsigned long long seed;
    const long long* fixed_idx;  // optional batch indices (tests)
};

struct TrainBuffers {
    long long* idx;
    float* z[3];
    __bf16* h[3];
    __bf16* ht[3];
    float* dz[4];
    __bf16* dzb[4];
    __bf16* dzt[4];
    __bf16* xt;
    float* grad;
    float* loss;  // kLossSlots per-step losses
    float* partial;  // split weight-gradient products, a slot each layer: [4][kGradSplits][128][128]
};

constexpr int kGradSplits = 32;  // workgroups along the batch of a weight-gradient product

constexpr int kLossSlots = 1024;

}  // namespace mlp

hipError_t mlp_refresh(const mlp::Model& md, hipStream_t s);
// One training step: sample (or take the fixed) batch, forward, backward into
// buf.grad, the batch loss into *loss (device), and an Adam update if asked.
hipError_t mlp_step(const mlp::Model& md, const mlp::TrainBatch& tb, const mlp::TrainBuffers& buf, int step,
                    bool update, float lr, float* loss, hipStream_t s);
// The prediction for one image, packed [outputs / 4][w*h][RGBA]; one fused
// kernel, no workspace.
hipError_t mlp_predict(const mlp::Model& md, const double* rgb, int w, int h, int outputs, double* out,
                       hipStream_t s);

}  // namespace ubhip
