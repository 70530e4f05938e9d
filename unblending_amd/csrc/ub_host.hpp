// ub_host.hpp -- host-side pieces of the HIP backend shared by the C interface
// (ub_capi.cpp) and the device pipeline (ub_pipeline.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "unblending.h"
#include "unblending_hip.h"
#include "ub_kernels.hpp"
#include "ub_mlp.hpp"

namespace ubhip {

// A device failure; maps to UNBLENDING_DEVICE_ERROR as the reference's
// gpu::ExecutionError maps to UNBLENDING_CUDA_ERROR (unblending_c.cpp:88).
struct ExecutionError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

extern thread_local std::string g_last_error;

// A validated layer set.
struct LayerSet {
    std::vector<unblending_layer> host;
    bool all_plus = false, all_so = false;
};

LayerSet make_layer_set(const unblending_layer* layers, size_t count);
size_t validate_pixels(unblending_rgb_view& image, size_t layer_count);

// One device-resident pass (first pass when first == nullptr).
struct DeviceJob {
    int w = 0, h = 0;
    const LayerSet* set = nullptr;
    const double* rgb = nullptr;       // [n][3]
    const double* first = nullptr;     // packed first pass (refinement)
    const double* alpha_in = nullptr;  // optional alpha planes of the first pass
    double* out = nullptr;             // packed output
    double* alpha_out = nullptr;       // optional alpha planes of the output (first pass)
    bool opaque = true, smooth = false;
    unblending_hip_params params{};
    void* workspace = nullptr;
    size_t workspace_size = 0;
    hipStream_t stream = nullptr;
    unblending_hip_stats* stats = nullptr;
};

// (s, y) pairs the inner minimiser keeps for n variables: memory when > 0,
// else NLopt's default vector storage for LD_LBFGS (the reference never sets
// it): max(1310720 / n, 10), lowered to max_evals when max_evals is not above
// that (NLopt src/algs/luksan/plis.c) -- 1000 at the reference's maxeval.
int lbfgs_memory(int n, int max_evals, int memory);
// Throws std::invalid_argument for problem variants the device does not solve.
void check_variants(const unblending_hip_params& p, int layers);
constexpr int kMaxEvals = 1000;  // unblending.cpp:249

// Copies a caller's params (struct_size bytes of them, at least up to
// warm_outer_cap; later fields keep their defaults).  False when too short.
bool read_params(const unblending_hip_params* params, unblending_hip_params& p);

size_t workspace_bytes(size_t n, size_t layers, bool refine, bool smooth, const unblending_hip_params& params);
size_t mlp_train_bytes(int batch);
mlp::TrainBuffers mlp_train_view(void* ws, int batch);
size_t filter_workspace_bytes(size_t n, int channels);
void run_device(const DeviceJob& job);
void run_filter(const double* rgb, int w, int h, const Channel* in, double* const* q, int channels, int radius,
                double eps, double* ab, hipStream_t stream, unblending_hip_stats* stats, bool timing, int block = 0,
                int rows = 0);

// Batches of device-resident images (unblending_hip_*_batch_device).
size_t batch_workspace_bytes(const unblending_hip_image* images, size_t count, bool refine, bool smooth,
                             const unblending_hip_params& params);
void run_batch(const unblending_hip_image* images, const std::vector<LayerSet>& sets, bool refine, bool opaque,
               bool smooth, const unblending_hip_params& params, void* workspace, hipStream_t stream,
               unblending_hip_stats* stats);
void host_solve(const unblending_rgb_view& rgb, const LayerSet& set, const double* first, bool opaque,
                bool smooth, const unblending_hip_params& params, double* output, unblending_hip_stats* stats);
// frees the device buffers host_solve keeps between calls on the current device
void release_host_buffers();

void run_composite(const double* layers, int w, int h, size_t count, const int32_t* comp_ops, const int32_t* modes,
                   double* out, hipStream_t stream);
void host_composite(const double* layers, int w, int h, size_t count, const int32_t* comp_ops, const int32_t* modes,
                    double* out);

hipError_t launch_zero(unsigned long long* p, int count, hipStream_t stream);

// The allocation-free filter entry's device scratch (ub_noalloc.hip): its size,
// and its device address on the current device (looked up once per device).
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
hipError_t noalloc_scratch(void** ptr);
hipError_t launch_store_desc(ImageDesc* dst, const ImageDesc& d, hipStream_t stream);

}  // namespace ubhip
