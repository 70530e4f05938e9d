// ub_pipeline.cpp -- the device pipeline of both passes.
//
// First pass (unblending/src/unblending.cpp:441-475): one solve launch.
// Refinement (unblending.cpp:335-439): guided filter of every alpha (and the
// background colour when smoothing) in groups of up to kMaxGroup channels --
// pass A then pass B per group -- into filtered planes that the re-solve
// launch reads directly (crop and normalisation happen in its prologue).
// The host-pointer path (host_solve) adds the uploads and the download, as
// the reference's solve_cuda does (pixel_pipeline.cu:63-105), on a stream of
// its own, and returns after synchronising it.
#include <algorithm>
#include <cstring>
#include <memory>

#include "ub_host.hpp"

namespace ubhip {
namespace {

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw ExecutionError(std::string(what) + ": " + hipGetErrorString(e));
}

constexpr size_t kHeader = 256;  // counters

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

int filter_channels(size_t layers, bool smooth) { return int(layers) + (smooth ? 3 : 0); }

DevLayer to_device(const unblending_layer& l)
{
    DevLayer d{};
    d.X = double(l.comp_x);
    d.Y = double(l.comp_y);
    d.Z = double(l.comp_z);
    for (int c = 0; c < 3; ++c) d.mean[c] = l.mean[c];
    for (int k = 0; k < 9; ++k) d.P[k] = l.precision[k];
    d.mode = l.blend_mode;
    return d;
}

SolverParams default_solver(const unblending_hip_params& p)
{
    SolverParams s;
    s.max_evals = 1000;   // unblending.cpp:249
    s.max_count = 20;     // :246
    s.max_ls = 20;
    s.alpha_fixed = p.alpha_fixed ? 1 : 0;
    s.ftol = 5e-3;        // :222
    s.xtol = 5e-3;
    s.eps = 5e-3;         // :221
    s.gamma = 0.25;       // :220
    s.beta = 10.0;        // :223
    s.rho0 = 100.0;       // :224
    return s;
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    explicit Events(bool on)
    {
        if (on) {
            check(hipEventCreate(&a), "hipEventCreate");
            check(hipEventCreate(&b), "hipEventCreate");
        }
    }
    ~Events()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start(hipStream_t s) { if (a) check(hipEventRecord(a, s), "hipEventRecord"); }
    void stop(hipStream_t s) { if (b) check(hipEventRecord(b, s), "hipEventRecord"); }
    double ms()
    {
        if (!a) return 0.0;
        check(hipEventSynchronize(b), "hipEventSynchronize");
        float v = 0.f;
        check(hipEventElapsedTime(&v, a, b), "hipEventElapsedTime");
        return double(v);
    }
};

struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes)
    {
        if (bytes) {
            const hipError_t e = hipMalloc(&p, bytes);
            if (e == hipErrorOutOfMemory) throw std::bad_alloc();
            check(e, "hipMalloc");
        }
    }
    ~DeviceBuffer() { if (p) (void)hipFree(p); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

void require_device()
{
    int devices = 0;
    const hipError_t e = hipGetDeviceCount(&devices);
    if (e != hipSuccess || devices == 0) throw ExecutionError("No HIP device is available");
}

}  // namespace

size_t filter_workspace_bytes(size_t n, int channels)
{
    const size_t group = size_t(std::min(channels, kMaxGroup));
    return align256(4 * group * n * sizeof(double));
}

namespace {
constexpr long long kPredictChunk = 1 << 18;  // pixels per prediction chunk

// Carves `bytes` (256-aligned) from a workspace cursor.
template <class T>
T* carve(char*& cur, size_t count)
{
    T* p = reinterpret_cast<T*>(cur);
    cur += align256(count * sizeof(T));
    return p;
}
}  // namespace

namespace {
// rows of one prediction chunk for an n-pixel image (a multiple of 32)
long long predict_chunk(size_t n) { return std::min<long long>(kPredictChunk, (long long)((n + 31) / 32 * 32)); }
}  // namespace

size_t mlp_predict_bytes(size_t n) { return 3 * align256(size_t(predict_chunk(n)) * mlp::kHidden * 2); }

mlp::PredictBuffers mlp_predict_view(void* ws, size_t n)
{
    mlp::PredictBuffers b;
    char* cur = static_cast<char*>(ws);
    b.chunk = predict_chunk(n);
    for (int l = 0; l < 3; ++l) b.h[l] = carve<__bf16>(cur, size_t(b.chunk) * mlp::kHidden);
    return b;
}

mlp::TrainBuffers mlp_train_view(void* ws, int batch)
{
    using namespace mlp;
    mlp::TrainBuffers b;
    char* cur = static_cast<char*>(ws);
    const size_t B = size_t(batch);
    b.idx = carve<long long>(cur, B);
    for (int l = 0; l < 3; ++l) {
        b.z[l] = carve<float>(cur, B * kHidden);
        b.h[l] = carve<__bf16>(cur, B * kHidden);
        b.ht[l] = carve<__bf16>(cur, B * kHidden);
    }
    for (int l = 0; l < 4; ++l) {
        const size_t N = (l == 3) ? kOutPad : kHidden;
        b.dz[l] = carve<float>(cur, B * N);
        b.dzb[l] = carve<__bf16>(cur, B * N);
        b.dzt[l] = carve<__bf16>(cur, B * N);
    }
    b.xt = carve<__bf16>(cur, B * kInPad);
    b.grad = carve<float>(cur, kParamsPadded);
    b.loss = carve<float>(cur, mlp::kLossSlots);
    return b;
}

size_t mlp_train_bytes(int batch)
{
    // a dry run of the carving over a null base gives the size
    const mlp::TrainBuffers b = mlp_train_view(nullptr, batch);
    return size_t(reinterpret_cast<char*>(b.loss) - static_cast<char*>(nullptr)) + align256(mlp::kLossSlots * sizeof(float));
}

size_t workspace_bytes(size_t n, size_t layers, bool refine, bool smooth, bool warm)
{
    size_t bytes = kHeader + align256(size_t(kLaneStateDoubles) * sizeof(double));
    if (warm && !refine) bytes += mlp_predict_bytes(n) + align256(n * sizeof(long long));
    if (refine) {
        const int C = filter_channels(layers, smooth);
        bytes += filter_workspace_bytes(n, C);
        bytes += align256(size_t(C) * n * sizeof(double));
    }
    return bytes;
}

void run_filter(const double* rgb, int w, int h, const Channel* in, double* const* q, int channels, int radius,
                double eps, double* ab, hipStream_t stream, unblending_hip_stats* stats, bool timing)
{
    if (radius > kMaxFilterRadius)
        throw ExecutionError("Guided-filter radius is outside the device filter's domain");
    Events ea(timing), eb(timing);
    double coeff_ms = 0.0, out_ms = 0.0;
    int launches = 0;
    for (int g0 = 0; g0 < channels; g0 += kMaxGroup) {
        FilterArgs fa{};
        fa.rgb = rgb;
        fa.w = w;
        fa.h = h;
        fa.r = radius;
        fa.eps = eps;
        fa.groups = std::min(kMaxGroup, channels - g0);
        for (int j = 0; j < fa.groups; ++j) {
            fa.in[j] = in[g0 + j];
            fa.q[j] = q[g0 + j];
        }
        fa.ab = ab;
        ea.start(stream);
        check(launch_filter_coefficients(fa, stream), "guided filter (coefficients)");
        ea.stop(stream);
        eb.start(stream);
        check(launch_filter_output(fa, stream), "guided filter (output)");
        eb.stop(stream);
        launches += 2;
        if (timing) {
            coeff_ms += ea.ms();
            out_ms += eb.ms();
        }
    }
    if (stats) {
        stats->filter_coeff_ms = coeff_ms;
        stats->filter_out_ms = out_ms;
        stats->filter_ms = coeff_ms + out_ms;
        stats->filter_radius = radius;
        stats->filter_channels = channels;
        stats->filter_launches = launches;
    }
}

void run_device(const DeviceJob& job)
{
    const size_t L = job.set->host.size();
    if (L < 1 || L > size_t(kMaxLayers))
        throw ExecutionError("Layer count is outside the device solver's domain (1 to 8 layers)");
    const size_t n = size_t(job.w) * size_t(job.h);
    const bool refine = job.first != nullptr;
    const bool timing = job.params.timing != 0;
    char* ws = static_cast<char*>(job.workspace);
    auto* counters = reinterpret_cast<unsigned long long*>(ws);

    PixelArgs args{};
    for (size_t l = 0; l < L; ++l) args.layers[l] = to_device(job.set->host[l]);
    args.rgb = job.rgb;
    args.n = (long long)n;
    args.sp = default_solver(job.params);
    args.opaque = job.opaque ? 1 : 0;
    args.smooth = job.smooth ? 1 : 0;
    args.all_plus = job.set->all_plus ? 1 : 0;
    args.all_so = job.set->all_so ? 1 : 0;
    args.solver = job.params.solver;
    args.out = job.out;
    args.alpha_planes = job.alpha_out;
    args.counters = counters;
    args.lane_state = reinterpret_cast<double*>(ws + kHeader);
    check(launch_zero(counters, kNumCounters, job.stream), "zero counters");

    if (job.stats) *job.stats = unblending_hip_stats{};
    const bool warm = !refine && job.params.warm_start != 0;
    double predict_ms = 0.0;
    if (warm) {
        if (!job.params.mlp) throw std::invalid_argument("A warm start needs a model");
        if (4 * L > size_t(mlp::kOutPad)) throw ExecutionError("The warm-start model predicts at most 8 layers");
        char* wcur = ws + kHeader + align256(size_t(kLaneStateDoubles) * sizeof(double));
        const mlp::PredictBuffers pb = mlp_predict_view(wcur, n);
        args.fallback = reinterpret_cast<long long*>(wcur + mlp_predict_bytes(n));
        const mlp::Model md = mlp::model_view(const_cast<void*>(job.params.mlp));
        Events ep(timing);
        ep.start(job.stream);
        // the prediction is written into the output buffer, which the solve
        // reads pixel by pixel before writing the same pixel's solution
        check(mlp_predict(md, job.rgb, job.w, job.h, int(4 * L), job.out, pb, job.stream), "warm-start prediction");
        ep.stop(job.stream);
        predict_ms = ep.ms();
        args.warm = job.out;
    }
    args.queue = kQueue;
    if (refine) {
        int radius = job.params.filter_radius;
        if (radius < 0) radius = 60 * std::min(job.w, job.h) / 1000;  // unblending.cpp:351
        const double eps = job.params.filter_epsilon > 0.0 ? job.params.filter_epsilon : 1e-4;  // :353
        const int C = filter_channels(L, job.smooth);
        const size_t base = kHeader + align256(size_t(kLaneStateDoubles) * sizeof(double));
        double* ab = reinterpret_cast<double*>(ws + base);
        double* qbase = reinterpret_cast<double*>(ws + base + filter_workspace_bytes(n, C));
        std::vector<Channel> in(C);
        std::vector<double*> q(C);
        for (size_t l = 0; l < L; ++l)
            in[l] = job.alpha_in ? Channel{job.alpha_in + l * n, 1} : Channel{job.first + 4 * l * n + 3, 4};
        for (int c = 0; c < (job.smooth ? 3 : 0); ++c) in[L + c] = Channel{job.first + c, 4};
        for (int c = 0; c < C; ++c) q[c] = qbase + size_t(c) * n;
        run_filter(job.rgb, job.w, job.h, in.data(), q.data(), C, radius, eps, ab, job.stream, job.stats, timing);
        args.first = job.first;
        args.filtered = qbase;
    }

    Events es(timing);
    es.start(job.stream);
    if (warm && args.solver != 0) throw std::invalid_argument("The warm start runs on the persistent solve kernel");
    check(launch_solve(int(L), refine, args, job.stream), "solve");
    if (warm) {
        // re-solve, from the default initialisation, the pixels whose
        // warm-started solve missed the tolerances
        PixelArgs fb = args;
        fb.warm = nullptr;
        fb.list = args.fallback;
        fb.fallback = nullptr;
        fb.queue = kQueue2;
        check(launch_solve(int(L), false, fb, job.stream), "fallback solve");
    }
    es.stop(job.stream);
    if (job.stats) {
        unsigned long long host_counters[kNumCounters] = {};
        check(hipMemcpyAsync(host_counters, counters, sizeof host_counters, hipMemcpyDeviceToHost, job.stream),
              "counters");
        check(hipStreamSynchronize(job.stream), "hipStreamSynchronize");
        job.stats->evals = host_counters[kEvals];
        job.stats->iterations = host_counters[kIterations];
        job.stats->outer = host_counters[kOuter];
        job.stats->pixels = host_counters[kPixels];
        job.stats->fallback_pixels = host_counters[kFallback];
        job.stats->solve_ms = es.ms();
        job.stats->predict_ms = predict_ms;
    }
}

void host_solve(const unblending_rgb_view& rgb, const LayerSet& set, const double* first, bool opaque,
                bool smooth, const unblending_hip_params& params, double* output, unblending_hip_stats* stats)
{
    require_device();
    const size_t w = size_t(rgb.width), h = size_t(rgb.height), n = w * h, L = set.host.size();
    if (L > size_t(kMaxLayers))
        throw ExecutionError("Layer count is outside the device solver's domain (1 to 8 layers)");
    Stream stream;
    DeviceBuffer d_rgb(3 * n * sizeof(double));
    DeviceBuffer d_out(4 * L * n * sizeof(double));
    DeviceBuffer d_first(first ? 4 * L * n * sizeof(double) : 0);
    const bool warm = first == nullptr && params.warm_start != 0;
    const size_t ws_bytes = workspace_bytes(n, L, first != nullptr, smooth, warm);
    DeviceBuffer d_ws(ws_bytes);
    check(hipMemcpy2DAsync(d_rgb.p, 3 * w * sizeof(double), rgb.data, rgb.row_stride * sizeof(double),
                           3 * w * sizeof(double), h, hipMemcpyHostToDevice, stream.s),
          "upload rgb");
    if (first)
        check(hipMemcpyAsync(d_first.p, first, 4 * L * n * sizeof(double), hipMemcpyHostToDevice, stream.s),
              "upload first pass");
    DeviceJob job;
    job.w = int(w);
    job.h = int(h);
    job.set = &set;
    job.rgb = d_rgb.as<double>();
    job.first = first ? d_first.as<double>() : nullptr;
    job.out = d_out.as<double>();
    job.opaque = opaque;
    job.smooth = smooth;
    job.params = params;
    job.workspace = d_ws.p;
    job.workspace_size = ws_bytes;
    job.stream = stream.s;
    job.stats = stats;
    run_device(job);
    check(hipStreamSynchronize(stream.s), "solve");
    check(hipMemcpyAsync(output, d_out.p, 4 * L * n * sizeof(double), hipMemcpyDeviceToHost, stream.s),
          "download");
    check(hipStreamSynchronize(stream.s), "download");
}

}  // namespace ubhip
