// ub_capi.cpp -- the exported C interface of the HIP backend.
//
// Argument checks follow the reference's C ABI and packed-buffer API in the
// same order (unblending/sdk/unblending_c.cpp:23-92 and
// unblending/src/unblending.cpp:508-534, 590-614), with the same status codes
// (unblending_c.cpp:87-91).  The solve itself runs on the current HIP device
// through ub_pipeline.cpp; there is no CPU solver in this library.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "ub_host.hpp"

namespace ubhip {

thread_local std::string g_last_error;

namespace {

enum class Backend { automatic, cpu, cuda, hip };

// f(begin, end) over [0, n) in up to 8 ranges of at least 2^18 items on host
// threads; with fewer threads available the calling thread runs the rest.
template <class F>
void parallel_for(size_t n, F f)
{
    constexpr size_t kMinRange = size_t(1) << 18;
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t threads = std::max<size_t>(1, std::min<size_t>(hw ? hw : 1, 8));
    const size_t parts = std::min(threads, (n + kMinRange - 1) / kMinRange);
    if (parts <= 1) {
        f(size_t(0), n);
        return;
    }
    const size_t step = (n + parts - 1) / parts;
    auto range = [&](size_t p) {
        const size_t b = std::min(n, p * step);
        f(b, std::min(n, b + step));
    };
    std::vector<std::thread> pool;
    pool.reserve(parts - 1);
    size_t started = 1;  // range 0 is the calling thread's
    try {
        for (; started < parts; ++started) pool.emplace_back(range, started);
    } catch (...) {
        // no more threads: the ranges not started run below
    }
    range(0);
    for (size_t p = started; p < parts; ++p) range(p);
    for (auto& t : pool) t.join();
}

// unblending_c.cpp:12-21, plus the HIP value.
Backend select_backend(int32_t backend)
{
    switch (backend) {
    case UNBLENDING_BACKEND_AUTO: return Backend::automatic;
    case UNBLENDING_BACKEND_CPU: return Backend::cpu;
    case UNBLENDING_BACKEND_CUDA: return Backend::cuda;
    case UNBLENDING_BACKEND_HIP: return Backend::hip;
    default: throw std::invalid_argument("Unknown backend");
    }
}

// Eigen's isApprox for P against P^T: |P - P^T|^2 <= prec^2 min(|P|^2, |P^T|^2).
bool symmetric_approx(const double* P)
{
    double diff = 0.0, norm = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double d = P[3 * r + c] - P[3 * c + r];
            diff += d * d;
            norm += P[3 * r + c] * P[3 * r + c];
        }
    return diff <= 1e-12 * 1e-12 * norm;
}

// Cholesky of the lower triangle; fails on a non-positive (or NaN) pivot as
// Eigen::LLT reports (unblending_c.cpp:42).
bool cholesky_ok(const double* P)
{
    double Lm[3][3] = {{0}};
    for (int k = 0; k < 3; ++k) {
        double x = P[3 * k + k];
        for (int j = 0; j < k; ++j) x -= Lm[k][j] * Lm[k][j];
        if (!(x > 0.0)) return false;
        Lm[k][k] = std::sqrt(x);
        for (int i = k + 1; i < 3; ++i) {
            double y = P[3 * i + k];
            for (int j = 0; j < k; ++j) y -= Lm[i][j] * Lm[k][j];
            Lm[i][k] = y / Lm[k][k];
        }
    }
    return true;
}

}  // namespace

// unblending_c.cpp:23-48
LayerSet make_layer_set(const unblending_layer* layers, size_t count)
{
    LayerSet set;
    set.host.assign(layers, layers + count);
    set.all_plus = true;
    set.all_so = true;
    for (size_t i = 0; i < count; ++i) {
        const unblending_layer& l = layers[i];
        if (l.blend_mode < UNBLENDING_NORMAL || l.blend_mode > UNBLENDING_LINEAR_DODGE)
            throw std::invalid_argument("Unknown blend mode");
        bool finite = true;
        for (double v : l.mean) finite = finite && std::isfinite(v);
        for (double v : l.precision) finite = finite && std::isfinite(v);
        if (!finite || !symmetric_approx(l.precision) || !cholesky_ok(l.precision))
            throw std::invalid_argument("Invalid Gaussian color model");
        set.all_plus = set.all_plus && l.comp_x == 2 && l.comp_y == 1 && l.comp_z == 1;
        set.all_so = set.all_so && l.comp_x == 1 && l.comp_y == 1 && l.comp_z == 1;
    }
    return set;
}

// unblending.cpp:508-534: normalises the stride and returns the packed size.
size_t validate_pixels(unblending_rgb_view& image, size_t layer_count)
{
    constexpr auto max_int = std::numeric_limits<int>::max();
    if (!image.data || image.width <= 0 || image.height <= 0 || image.width > max_int / image.height)
        throw std::invalid_argument("Invalid RGB image dimensions or data");
    const size_t max_doubles = std::vector<double>().max_size();
    if (size_t(image.width) > max_doubles / 3) throw std::length_error("RGB row is too large");
    const size_t row_size = 3 * size_t(image.width);
    if (!image.row_stride) image.row_stride = row_size;
    if (image.row_stride < row_size || image.size < row_size ||
        size_t(image.height - 1) > (image.size - row_size) / image.row_stride)
        throw std::invalid_argument("RGB buffer is too small or has an invalid row stride");
    if (layer_count == 0) throw std::invalid_argument("Each layer requires a color model");
    const size_t pixels = size_t(image.width) * size_t(image.height);
    if (layer_count > size_t(max_int) / 4 || pixels > max_doubles / 4 / layer_count)
        throw std::length_error("RGBA layer stack is too large");
    return 4 * pixels * layer_count;
}

namespace {

struct FirstPass {
    const double* data;
    size_t size;
};

// unblending_c.cpp:50-92 with the HIP dispatch of gpu_dispatch.cpp:52-123.
int32_t solve(const unblending_rgb_view* image, const unblending_layer* layers, size_t layer_count,
              const unblending_options* options, const FirstPass* first,
              const unblending_hip_params* params, double* output, size_t output_capacity,
              unblending_hip_stats* stats)
{
    g_last_error.clear();
    try {
        constexpr auto max_int = std::numeric_limits<int>::max();
        if (!image || !layers || !layer_count || layer_count > size_t(max_int) / 4 || image->width <= 0 ||
            image->height <= 0 || image->width > max_int / image->height)
            return UNBLENDING_INVALID_ARGUMENT;
        const size_t pixels = size_t(image->width) * size_t(image->height);
        if (pixels > std::vector<double>().max_size() / 4 / layer_count) return UNBLENDING_INVALID_ARGUMENT;
        const size_t required = 4 * pixels * layer_count;
        if (output_capacity < required) return UNBLENDING_BUFFER_TOO_SMALL;
        if (!output) return UNBLENDING_INVALID_ARGUMENT;

        const unblending_options defaults = UNBLENDING_OPTIONS_INIT;
        if (!options) options = &defaults;
        if (options->struct_size < sizeof(unblending_options) ||
            (options->opaque_background != 0 && options->opaque_background != 1) ||
            (options->smooth_background != 0 && options->smooth_background != 1) ||
            options->target_concurrency < 0)
            return UNBLENDING_INVALID_ARGUMENT;
        unblending_hip_params p;
        if (!read_params(params, p)) return UNBLENDING_INVALID_ARGUMENT;

        if (p.lbfgs_memory < 0) return UNBLENDING_INVALID_ARGUMENT;
        const Backend backend = select_backend(options->backend);
        const LayerSet set = make_layer_set(layers, layer_count);
        const bool opaque = options->opaque_background != 0;
        const bool smooth = options->smooth_background != 0;

        unblending_rgb_view rgb = *image;
        const size_t size = validate_pixels(rgb, layer_count);
        if (first) {
            // unblending.cpp:597-604
            if (!first->data || first->size != size)
                throw std::invalid_argument("Refinement requires one packed RGBA image per layer");
            if (smooth && !opaque) throw std::invalid_argument("Smoothing requires an opaque background");
            if (!set.all_plus && !set.all_so)
                throw std::invalid_argument("Refinement requires an all-Plus or all-SourceOver stack");
        }
        if (backend == Backend::cpu)
            throw std::invalid_argument("The CPU backend is not part of the HIP library");
        if (backend == Backend::cuda) throw ExecutionError("This library is built without CUDA");

        host_solve(rgb, set, first ? first->data : nullptr, opaque, smooth && first != nullptr, p, output, stats);
        return UNBLENDING_OK;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of memory";
        return UNBLENDING_OUT_OF_MEMORY;
    } catch (const ExecutionError& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return UNBLENDING_INVALID_ARGUMENT;
    } catch (const std::length_error& e) {
        g_last_error = e.what();
        return UNBLENDING_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    } catch (...) {
        return UNBLENDING_SOLVER_ERROR;
    }
}

// Checks shared by the device-resident calls.
int32_t device_call(const double* rgb, int32_t width, int32_t height, const unblending_layer* layers,
                    size_t layer_count, const double* first, const double* alpha_planes,
                    const unblending_options* options, const unblending_hip_params* params,
                    double* output, double* alpha_out, void* workspace, size_t workspace_size,
                    void* stream, unblending_hip_stats* stats)
{
    g_last_error.clear();
    try {
        constexpr auto max_int = std::numeric_limits<int>::max();
        if (!rgb || !layers || !layer_count || layer_count > size_t(max_int) / 4 || width <= 0 ||
            height <= 0 || width > max_int / height || !output)
            return UNBLENDING_INVALID_ARGUMENT;
        const unblending_options defaults = UNBLENDING_OPTIONS_INIT;
        if (!options) options = &defaults;
        if (options->struct_size < sizeof(unblending_options) ||
            (options->opaque_background != 0 && options->opaque_background != 1) ||
            (options->smooth_background != 0 && options->smooth_background != 1) ||
            options->target_concurrency < 0)
            return UNBLENDING_INVALID_ARGUMENT;
        unblending_hip_params p;
        if (!read_params(params, p)) return UNBLENDING_INVALID_ARGUMENT;
        const Backend backend = select_backend(options->backend);
        if (backend == Backend::cpu) return UNBLENDING_INVALID_ARGUMENT;
        if (backend == Backend::cuda) return UNBLENDING_DEVICE_ERROR;
        const LayerSet set = make_layer_set(layers, layer_count);
        const bool opaque = options->opaque_background != 0;
        const bool smooth = options->smooth_background != 0;
        if (first) {
            if (smooth && !opaque) return UNBLENDING_INVALID_ARGUMENT;
            if (!set.all_plus && !set.all_so) return UNBLENDING_INVALID_ARGUMENT;
        }
        if (p.lbfgs_memory < 0) return UNBLENDING_INVALID_ARGUMENT;
        const size_t need = workspace_bytes(size_t(width) * size_t(height), layer_count, first != nullptr,
                                            smooth && first != nullptr, p);
        if (!workspace || workspace_size < need) return UNBLENDING_BUFFER_TOO_SMALL;
        DeviceJob job;
        job.w = width;
        job.h = height;
        job.set = &set;
        job.rgb = rgb;
        job.first = first;
        job.alpha_in = alpha_planes;
        job.out = output;
        job.alpha_out = alpha_out;
        job.opaque = opaque;
        job.smooth = smooth && first != nullptr;
        job.params = p;
        job.workspace = workspace;
        job.workspace_size = workspace_size;
        job.stream = static_cast<hipStream_t>(stream);
        job.stats = stats;
        run_device(job);
        return UNBLENDING_OK;
    } catch (const std::bad_alloc&) {
        return UNBLENDING_OUT_OF_MEMORY;
    } catch (const ExecutionError& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return UNBLENDING_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    } catch (...) {
        return UNBLENDING_SOLVER_ERROR;
    }
}

// Batches: the checks of device_call for every image, in image order.
int32_t batch_call(const unblending_hip_image* images, size_t count, const unblending_options* options,
                   const unblending_hip_params* params, bool refine, void* workspace, size_t workspace_size,
                   void* stream, unblending_hip_stats* stats)
{
    g_last_error.clear();
    try {
        constexpr auto max_int = std::numeric_limits<int>::max();
        if (!images || count == 0) return UNBLENDING_INVALID_ARGUMENT;
        const unblending_options defaults = UNBLENDING_OPTIONS_INIT;
        if (!options) options = &defaults;
        if (options->struct_size < sizeof(unblending_options) ||
            (options->opaque_background != 0 && options->opaque_background != 1) ||
            (options->smooth_background != 0 && options->smooth_background != 1) ||
            options->target_concurrency < 0)
            return UNBLENDING_INVALID_ARGUMENT;
        unblending_hip_params p;
        if (!read_params(params, p)) return UNBLENDING_INVALID_ARGUMENT;
        if (p.solver != 0 || p.chunk_size < 0 || p.lbfgs_memory < 0) return UNBLENDING_INVALID_ARGUMENT;
        const Backend backend = select_backend(options->backend);
        if (backend == Backend::cpu) return UNBLENDING_INVALID_ARGUMENT;
        if (backend == Backend::cuda) return UNBLENDING_DEVICE_ERROR;
        const bool opaque = options->opaque_background != 0;
        const bool smooth = refine && options->smooth_background != 0;
        if (smooth && !opaque) return UNBLENDING_INVALID_ARGUMENT;
        std::vector<LayerSet> sets;
        sets.reserve(count);
        for (size_t i = 0; i < count; ++i) {
            const unblending_hip_image& im = images[i];
            if (!im.rgb || !im.layers || !im.layer_count || im.layer_count > size_t(max_int) / 4 || im.width <= 0 ||
                im.height <= 0 || im.width > max_int / im.height || !im.output || (refine && !im.first_pass))
                return UNBLENDING_INVALID_ARGUMENT;
            sets.push_back(make_layer_set(im.layers, im.layer_count));
            if (refine && !sets.back().all_plus && !sets.back().all_so) return UNBLENDING_INVALID_ARGUMENT;
            if (!refine && p.warm_start && im.mlp && 4 * im.layer_count > size_t(mlp::kOutPad))
                throw ExecutionError("The warm-start model predicts at most 8 layers");
            if (im.layer_count > size_t(kMaxLayers))
                throw ExecutionError("Layer count is outside the device solver's domain (1 to 8 layers)");
        }
        if (!workspace || workspace_size < batch_workspace_bytes(images, count, refine, smooth, p))
            return UNBLENDING_BUFFER_TOO_SMALL;
        run_batch(images, sets, refine, opaque, smooth, p, workspace, static_cast<hipStream_t>(stream), stats);
        return UNBLENDING_OK;
    } catch (const std::bad_alloc&) {
        return UNBLENDING_OUT_OF_MEMORY;
    } catch (const ExecutionError& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return UNBLENDING_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    } catch (...) {
        return UNBLENDING_SOLVER_ERROR;
    }
}

}  // namespace
}  // namespace ubhip

using namespace ubhip;

extern "C" {

UNBLENDING_API uint32_t unblending_abi_version(void) { return UNBLENDING_ABI_VERSION; }

UNBLENDING_API int32_t unblending_unmix(const unblending_rgb_view* image, const unblending_layer* layers,
                                        size_t layer_count, const unblending_options* options,
                                        double* output, size_t output_capacity)
{
    return solve(image, layers, layer_count, options, nullptr, nullptr, output, output_capacity, nullptr);
}

UNBLENDING_API int32_t unblending_refine(const unblending_rgb_view* image, const unblending_layer* layers,
                                         size_t layer_count, const double* first_pass, size_t first_pass_size,
                                         const unblending_options* options, double* output,
                                         size_t output_capacity)
{
    const FirstPass first{first_pass, first_pass_size};
    return solve(image, layers, layer_count, options, &first, nullptr, output, output_capacity, nullptr);
}

UNBLENDING_API int32_t unblending_hip_unmix_ex(const unblending_rgb_view* image, const unblending_layer* layers,
                                               size_t layer_count, const unblending_options* options,
                                               const unblending_hip_params* params, double* output,
                                               size_t output_capacity, unblending_hip_stats* stats)
{
    return solve(image, layers, layer_count, options, nullptr, params, output, output_capacity, stats);
}

UNBLENDING_API int32_t unblending_hip_refine_ex(const unblending_rgb_view* image, const unblending_layer* layers,
                                                size_t layer_count, const double* first_pass,
                                                size_t first_pass_size, const unblending_options* options,
                                                const unblending_hip_params* params, double* output,
                                                size_t output_capacity, unblending_hip_stats* stats)
{
    const FirstPass first{first_pass, first_pass_size};
    return solve(image, layers, layer_count, options, &first, params, output, output_capacity, stats);
}

UNBLENDING_API size_t unblending_hip_workspace_size(int32_t width, int32_t height, size_t layer_count,
                                                   int32_t refine, int32_t smooth_background, int32_t warm_start)
{
    (void)warm_start;  // the warm start's prediction needs no workspace
    return unblending_hip_workspace_size_ex(width, height, layer_count, refine, smooth_background, nullptr);
}

UNBLENDING_API size_t unblending_hip_workspace_size_ex(int32_t width, int32_t height, size_t layer_count,
                                                      int32_t refine, int32_t smooth_background,
                                                      const unblending_hip_params* params)
{
    unblending_hip_params p;
    if (width <= 0 || height <= 0 || layer_count == 0 || !read_params(params, p) || p.lbfgs_memory < 0) return 0;
    return workspace_bytes(size_t(width) * size_t(height), layer_count, refine != 0, smooth_background != 0, p);
}

UNBLENDING_API int32_t unblending_hip_unmix_device(const double* rgb, int32_t width, int32_t height,
                                                   const unblending_layer* layers, size_t layer_count,
                                                   const unblending_options* options,
                                                   const unblending_hip_params* params, double* output,
                                                   double* alpha_planes, void* workspace, size_t workspace_size,
                                                   void* stream, unblending_hip_stats* stats)
{
    return device_call(rgb, width, height, layers, layer_count, nullptr, nullptr, options, params, output,
                       alpha_planes, workspace, workspace_size, stream, stats);
}

UNBLENDING_API int32_t unblending_hip_refine_device(const double* rgb, int32_t width, int32_t height,
                                                    const unblending_layer* layers, size_t layer_count,
                                                    const double* first_pass, const double* alpha_planes,
                                                    const unblending_options* options,
                                                    const unblending_hip_params* params, double* output,
                                                    void* workspace, size_t workspace_size, void* stream,
                                                    unblending_hip_stats* stats)
{
    if (!first_pass) return UNBLENDING_INVALID_ARGUMENT;
    return device_call(rgb, width, height, layers, layer_count, first_pass, alpha_planes, options, params,
                       output, nullptr, workspace, workspace_size, stream, stats);
}

UNBLENDING_API size_t unblending_hip_batch_workspace_size(const unblending_hip_image* images, size_t count,
                                                         int32_t refine, int32_t smooth_background)
{
    return unblending_hip_batch_workspace_size_ex(images, count, refine, smooth_background, nullptr);
}

UNBLENDING_API size_t unblending_hip_batch_workspace_size_ex(const unblending_hip_image* images, size_t count,
                                                            int32_t refine, int32_t smooth_background,
                                                            const unblending_hip_params* params)
{
    unblending_hip_params p;
    if (!images || count == 0 || !read_params(params, p) || p.lbfgs_memory < 0) return 0;
    for (size_t i = 0; i < count; ++i)
        if (images[i].width <= 0 || images[i].height <= 0 || images[i].layer_count == 0) return 0;
    return batch_workspace_bytes(images, count, refine != 0, refine != 0 && smooth_background != 0, p);
}

UNBLENDING_API int32_t unblending_hip_unmix_batch_device(const unblending_hip_image* images, size_t count,
                                                         const unblending_options* options,
                                                         const unblending_hip_params* params, void* workspace,
                                                         size_t workspace_size, void* stream,
                                                         unblending_hip_stats* stats)
{
    return batch_call(images, count, options, params, false, workspace, workspace_size, stream, stats);
}

UNBLENDING_API int32_t unblending_hip_refine_batch_device(const unblending_hip_image* images, size_t count,
                                                          const unblending_options* options,
                                                          const unblending_hip_params* params, void* workspace,
                                                          size_t workspace_size, void* stream,
                                                          unblending_hip_stats* stats)
{
    return batch_call(images, count, options, params, true, workspace, workspace_size, stream, stats);
}

// Filter launch shapes: 0 (by radius); radii up to 96 run one-wave strips (128
// or 256 columns; other widths mean 256), larger radii workgroup strips of
// 256, 512 or 1024 threads, wider than 2r.
static bool filter_block_ok(int32_t block, int32_t radius)
{
    if (block == 0) return true;
    if (block != 128 && block != 256 && block != 512 && block != 1024) return false;
    return radius <= ubhip::kWaveStripMaxRadius || (block != 128 && block - 2 * radius >= 1);
}

UNBLENDING_API size_t unblending_hip_filter_workspace_size(int32_t width, int32_t height, int32_t channels)
{
    if (width <= 0 || height <= 0 || channels <= 0) return 0;
    return filter_workspace_bytes(size_t(width) * size_t(height), channels);
}

UNBLENDING_API int32_t unblending_hip_guided_filter_device(const double* rgb, int32_t width, int32_t height,
                                                           const double* const* p, double* const* q,
                                                           int32_t channels, int32_t radius, double epsilon,
                                                           int32_t block, int32_t rows, void* workspace,
                                                           size_t workspace_size, void* stream,
                                                           unblending_hip_stats* stats)
{
    g_last_error.clear();
    try {
        if (!rgb || !p || !q || channels <= 0 || width <= 0 || height <= 0 || radius < 0 || !(epsilon > 0.0) ||
            !filter_block_ok(block, radius) || rows < 0)
            return UNBLENDING_INVALID_ARGUMENT;
        if (!workspace || workspace_size < filter_workspace_bytes(size_t(width) * size_t(height), channels))
            return UNBLENDING_BUFFER_TOO_SMALL;
        std::vector<Channel> in(channels);
        for (int c = 0; c < channels; ++c) in[c] = Channel{p[c], 1};
        run_filter(rgb, width, height, in.data(), q, channels, radius, epsilon, static_cast<double*>(workspace),
                   static_cast<hipStream_t>(stream), stats, stats != nullptr, block, rows);
        return UNBLENDING_OK;
    } catch (const ExecutionError& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    }
}

UNBLENDING_API int32_t unblending_hip_guided_filter_device_noalloc(const double* rgb, int32_t width, int32_t height,
                                                                   const double* const* p, double* const* q,
                                                                   int32_t channels, int32_t radius, double epsilon,
                                                                   int32_t block, int32_t rows, void* stream,
                                                                   unblending_hip_stats* stats)
{
    g_last_error.clear();
    try {
        if (!rgb || !p || !q || channels <= 0 || channels > kMaxChannels || width <= 0 || height <= 0 ||
            width > std::numeric_limits<int>::max() / height || radius < 0 || !(epsilon > 0.0) ||
            !filter_block_ok(block, radius) || rows < 0)
            return UNBLENDING_INVALID_ARGUMENT;
        const size_t need = filter_workspace_bytes(size_t(width) * size_t(height), channels);
        if (need > kNoAllocScratchBytes) {
            g_last_error = "the filter's scratch (" + std::to_string(need) + " bytes for " +
                           std::to_string(channels) + " channels of " + std::to_string(width) + " x " +
                           std::to_string(height) + " pixels) exceeds the allocation-free entry's " +
                           std::to_string(kNoAllocScratchBytes) + " bytes";
            return UNBLENDING_INVALID_ARGUMENT;
        }
        void* ws = nullptr;
        const hipError_t e = noalloc_scratch(&ws);
        if (e != hipSuccess) throw ExecutionError(std::string("scratch address: ") + hipGetErrorString(e));
        Channel in[kMaxChannels];
        for (int c = 0; c < channels; ++c) in[c] = Channel{p[c], 1};
        run_filter(rgb, width, height, in, q, channels, radius, epsilon, static_cast<double*>(ws),
                   static_cast<hipStream_t>(stream), stats, false, block, rows);
        return UNBLENDING_OK;
    } catch (const ExecutionError& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    }
}

// Arguments of the recomposite entries: dimensions, 1 to 16 layers, comp-op
// factors and blend modes of the reference's enums; the sizes in doubles.
static int32_t composite_args_ok(const double* layers, int32_t width, int32_t height, size_t count,
                                 const int32_t* comp_ops, const int32_t* modes, const double* output)
{
    if (!layers || !output || !comp_ops || !modes || width <= 0 || height <= 0 || count < 1 ||
        count > size_t(kMaxCompositeLayers) || width > std::numeric_limits<int>::max() / height)
        return UNBLENDING_INVALID_ARGUMENT;
    for (size_t k = 0; k < count; ++k)
        if (modes[k] < UNBLENDING_NORMAL || modes[k] > UNBLENDING_LINEAR_DODGE) return UNBLENDING_INVALID_ARGUMENT;
    return UNBLENDING_OK;
}

UNBLENDING_API int32_t unblending_hip_composite_device(const double* layers, int32_t width, int32_t height,
                                                       size_t layer_count, const int32_t* comp_ops,
                                                       const int32_t* modes, double* output, void* stream)
{
    g_last_error.clear();
    const int32_t st = composite_args_ok(layers, width, height, layer_count, comp_ops, modes, output);
    if (st != UNBLENDING_OK) return st;
    try {
        run_composite(layers, width, height, layer_count, comp_ops, modes, output, static_cast<hipStream_t>(stream));
        return UNBLENDING_OK;
    } catch (const ExecutionError& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    }
}

UNBLENDING_API int32_t unblending_hip_composite(const double* layers, size_t layers_size, int32_t width,
                                                int32_t height, size_t layer_count, const int32_t* comp_ops,
                                                const int32_t* modes, double* output, size_t output_capacity)
{
    g_last_error.clear();
    const int32_t st = composite_args_ok(layers, width, height, layer_count, comp_ops, modes, output);
    if (st != UNBLENDING_OK) return st;
    const size_t n = size_t(width) * size_t(height);
    if (layers_size != 4 * n * layer_count) return UNBLENDING_INVALID_ARGUMENT;
    if (output_capacity < 4 * n) return UNBLENDING_BUFFER_TOO_SMALL;
    try {
        host_composite(layers, width, height, layer_count, comp_ops, modes, output);
        return UNBLENDING_OK;
    } catch (const std::bad_alloc&) {
        return UNBLENDING_OUT_OF_MEMORY;
    } catch (const ExecutionError& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    }
}

UNBLENDING_API int32_t uhip_solve_planes(int32_t width, int32_t height, const unblending_layer* layers,
                                         size_t layer_count, const double* rgb_planes, const double* first_planes,
                                         int32_t opaque, int32_t smooth, double* out_planes)
{
    g_last_error.clear();
    try {
        // the checks of unblending_c.cpp:56-63 before anything is allocated
        constexpr auto max_int = std::numeric_limits<int>::max();
        if (width <= 0 || height <= 0 || width > max_int / height || !layers || !layer_count ||
            layer_count > size_t(max_int) / 4 || !rgb_planes || !out_planes)
            return UNBLENDING_INVALID_ARGUMENT;
        const size_t n = size_t(width) * size_t(height), L = layer_count;
        if (n > std::vector<double>().max_size() / 4 / L) return UNBLENDING_INVALID_ARGUMENT;
        // gpu_dispatch.cpp:136-158 planes -> packed, and back (:182-191).
        // the intermediate packed buffers are written in full before they are
        // read: allocated without zero-filling (3.8 GB each at 4096^2, L = 7)
        std::unique_ptr<double[]> rgb(new double[3 * n]), first, out(new double[4 * L * n]);
        parallel_for(n, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; ++i)
                for (int c = 0; c < 3; ++c) rgb[3 * i + c] = rgb_planes[c * n + i];
        });
        if (first_planes) {
            first.reset(new double[4 * L * n]);
            parallel_for(n, [&](size_t b, size_t e) {
                for (size_t l = 0; l < L; ++l)
                    for (size_t i = b; i < e; ++i) {
                        double* px = first.get() + 4 * (l * n + i);
                        for (int c = 0; c < 3; ++c) px[c] = first_planes[(L + 3 * l + c) * n + i];
                        px[3] = first_planes[l * n + i];
                    }
            });
        }
        unblending_rgb_view view{rgb.get(), 3 * n, width, height, 0};
        unblending_options opt = UNBLENDING_OPTIONS_INIT;
        opt.backend = UNBLENDING_BACKEND_HIP;
        opt.opaque_background = opaque ? 1 : 0;
        opt.smooth_background = smooth ? 1 : 0;
        const int32_t status = first_planes
            ? unblending_refine(&view, layers, L, first.get(), 4 * L * n, &opt, out.get(), 4 * L * n)
            : unblending_unmix(&view, layers, L, &opt, out.get(), 4 * L * n);
        if (status != UNBLENDING_OK) return status;
        parallel_for(n, [&](size_t b, size_t e) {
            for (size_t l = 0; l < L; ++l)
                for (size_t i = b; i < e; ++i) {
                    const double* px = out.get() + 4 * (l * n + i);
                    for (int c = 0; c < 3; ++c) out_planes[(L + 3 * l + c) * n + i] = px[c];
                    out_planes[l * n + i] = px[3];
                }
        });
        return UNBLENDING_OK;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of memory";
        return UNBLENDING_OUT_OF_MEMORY;
    } catch (const std::length_error& e) {
        g_last_error = e.what();
        return UNBLENDING_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_SOLVER_ERROR;
    } catch (...) {
        return UNBLENDING_SOLVER_ERROR;
    }
}

UNBLENDING_API size_t unblending_hip_mlp_model_size(void) { return mlp::model_bytes(); }

UNBLENDING_API int32_t unblending_hip_mlp_param_count(void) { return mlp::kParamsPadded; }

UNBLENDING_API int32_t unblending_hip_mlp_set_params(void* model, const float* params, void* stream)
{
    g_last_error.clear();
    if (!model || !params) return UNBLENDING_INVALID_ARGUMENT;
    const mlp::Model md = mlp::model_view(model);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(md.w, params, sizeof(float) * mlp::kParamsPadded, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(md.m, 0, 2 * sizeof(float) * mlp::kParamsPadded, s);
    if (e == hipSuccess) e = mlp_refresh(md, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        g_last_error = hipGetErrorString(e);
        return UNBLENDING_DEVICE_ERROR;
    }
    return UNBLENDING_OK;
}

UNBLENDING_API int32_t unblending_hip_mlp_get_params(const void* model, float* params, void* stream)
{
    g_last_error.clear();
    if (!model || !params) return UNBLENDING_INVALID_ARGUMENT;
    const mlp::Model md = mlp::model_view(const_cast<void*>(model));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(params, md.w, sizeof(float) * mlp::kParamsPadded, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        g_last_error = hipGetErrorString(e);
        return UNBLENDING_DEVICE_ERROR;
    }
    return UNBLENDING_OK;
}

UNBLENDING_API size_t unblending_hip_mlp_train_workspace_size(int32_t batch)
{
    if (batch <= 0 || batch % 64) return 0;
    return mlp_train_bytes(batch);
}

UNBLENDING_API int32_t unblending_hip_mlp_train(void* model, const double* images, int32_t count, int32_t width,
                                                int32_t height, const double* targets, int32_t layers,
                                                int32_t steps, int32_t steps_done, int32_t batch,
                                                float learning_rate, uint64_t seed, const int64_t* fixed_samples,
                                                int32_t update, void* workspace, size_t workspace_size, void* stream,
                                                float* losses, float* gradients)
{
    g_last_error.clear();
    if (!model || !images || !targets || count <= 0 || width <= 0 || height <= 0 || layers <= 0 ||
        4 * layers > mlp::kOutPad || steps < 0 || steps_done < 0 || batch <= 0 || batch % 64)
        return UNBLENDING_INVALID_ARGUMENT;
    if (!workspace || workspace_size < mlp_train_bytes(batch)) return UNBLENDING_BUFFER_TOO_SMALL;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const mlp::Model md = mlp::model_view(model);
    const mlp::TrainBuffers buf = mlp_train_view(workspace, batch);
    mlp::TrainBatch tb;
    tb.images = images;
    tb.targets = targets;
    tb.w = width;
    tb.h = height;
    tb.layers = layers;
    tb.outputs = 4 * layers;
    tb.batch = batch;
    tb.pool = (long long)count * width * height;
    tb.seed = seed;
    tb.fixed_idx = reinterpret_cast<const long long*>(fixed_samples);
    hipError_t e = hipSuccess;
    for (int step = 0; step < steps && e == hipSuccess; ++step) {
        const int slot = step % mlp::kLossSlots;
        e = mlp_step(md, tb, buf, steps_done + step, update != 0, learning_rate, buf.loss + slot, s);
        // per-step losses come back a block of slots at a time
        if (e == hipSuccess && losses && (slot == mlp::kLossSlots - 1 || step == steps - 1))
            e = hipMemcpyAsync(losses + (step - slot), buf.loss, sizeof(float) * (slot + 1), hipMemcpyDeviceToHost,
                               s);
        if (e == hipSuccess && losses && slot == mlp::kLossSlots - 1) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess) {
        g_last_error = hipGetErrorString(e);
        return UNBLENDING_DEVICE_ERROR;
    }
    if (gradients) e = hipMemcpyAsync(gradients, buf.grad, sizeof(float) * mlp::kParamsPadded, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        g_last_error = hipGetErrorString(e);
        return UNBLENDING_DEVICE_ERROR;
    }
    return UNBLENDING_OK;
}

UNBLENDING_API size_t unblending_hip_mlp_predict_workspace_size(void) { return 0; }

UNBLENDING_API int32_t unblending_hip_mlp_predict(const void* model, const double* rgb, int32_t width,
                                                  int32_t height, int32_t layers, double* output, void* workspace,
                                                  size_t workspace_size, void* stream)
{
    g_last_error.clear();
    if (!model || !rgb || !output || width <= 0 || height <= 0 || layers <= 0 || 4 * layers > mlp::kOutPad)
        return UNBLENDING_INVALID_ARGUMENT;
    (void)workspace;  // the fused prediction kernel needs none
    (void)workspace_size;
    const hipError_t e = mlp_predict(mlp::model_view(const_cast<void*>(model)), rgb, width, height, 4 * layers,
                                     output, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        g_last_error = hipGetErrorString(e);
        return UNBLENDING_DEVICE_ERROR;
    }
    return UNBLENDING_OK;
}

UNBLENDING_API const char* unblending_hip_last_error(void) { return g_last_error.c_str(); }

UNBLENDING_API int32_t unblending_hip_release_buffers(void)
{
    g_last_error.clear();
    try {
        release_host_buffers();
        return UNBLENDING_OK;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return UNBLENDING_DEVICE_ERROR;
    }
}

}  // extern "C"
