// ub_pipeline.cpp -- the device pipeline of both passes.
//
// First pass (unblending/src/unblending.cpp:441-475): one solve launch.
// Refinement (unblending.cpp:335-439): guided filter of every alpha (and the
// background colour when smoothing) -- a bounds pass, pass A (coefficients of
// every channel), pass B (outputs of every channel) -- into filtered planes
// that the re-solve launch reads directly (crop and normalisation happen in
// its prologue).
// The host-pointer path (host_solve) adds the uploads and the download, as
// the reference's solve_cuda does (pixel_pipeline.cu:63-105), on a stream of
// its own, and returns after synchronising it; the packed planes travel
// through pinned staging buffers.
#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ub_host.hpp"

namespace ubhip {
namespace {

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw ExecutionError(std::string(what) + ": " + hipGetErrorString(e));
}

// workspace header: counters at 0, then the image descriptor of the solve
constexpr size_t kDescOffset = 256;
constexpr size_t kDescStride = (sizeof(ImageDesc) + 255) / 256 * 256;
constexpr size_t kHeader = kDescOffset + kDescStride;

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

constexpr int kDefaultWarmOuterCap = 8;

SolverParams default_solver(const unblending_hip_params& p, int layers, bool refine)
{
    SolverParams s;
    s.gray = p.gray_layers != 0 ? 1 : 0;
    s.sparse_sigma = (p.sparsity && !refine) ? 10.0 : 0.0;  // unblending.cpp:234-240
    s.warm_max_count = p.warm_outer_cap > 0 ? p.warm_outer_cap : kDefaultWarmOuterCap;
    s.max_evals = kMaxEvals;  // unblending.cpp:249
    s.memory = lbfgs_memory(4 * layers, s.max_evals, p.lbfgs_memory);
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

// memcpy of a large block by up to 8 host threads (parts of at least 4 MiB);
// with fewer threads available the calling thread copies the rest.
void parallel_copy(void* dst, const void* src, size_t bytes)
{
    constexpr size_t kMinPart = size_t(4) << 20;
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t threads = std::max<size_t>(1, std::min<size_t>(hw ? hw : 1, 8));
    const size_t parts = std::min(threads, (bytes + kMinPart - 1) / kMinPart);
    auto copy_part = [dst, src](size_t off, size_t len) {
        if (len) std::memcpy(static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, len);
    };
    if (parts <= 1) {
        copy_part(0, bytes);
        return;
    }
    const size_t step = (bytes + parts - 1) / parts;
    auto part_len = [&](size_t p) { return p * step < bytes ? std::min(step, bytes - p * step) : size_t(0); };
    std::vector<std::thread> pool;
    pool.reserve(parts - 1);
    size_t started = 1;  // part 0 is the calling thread's
    try {
        for (; started < parts; ++started) pool.emplace_back(copy_part, started * step, part_len(started));
    } catch (...) {
        // no more threads: the parts not started are copied below
    }
    copy_part(0, part_len(0));
    for (size_t p = started; p < parts; ++p) copy_part(p * step, part_len(p));
    for (auto& t : pool) t.join();
}

// Copies between the caller's (pageable) host buffers and the device through
// two pinned staging buffers: host threads fill or drain one while the DMA
// engine moves the other.  The runtime's own copies from pageable memory
// run at a fraction of the link's rate (measured on an MI355X host: 30 GB/s
// up and 18 GB/s down, 57 GB/s from pinned memory), and a download into a
// freshly allocated result also pays its page faults on one thread.  The
// pinned buffers serve any device; the two events that order their reuse
// are created for each copy on the current device.
class Staging {
public:
    static constexpr size_t kChunk = size_t(64) << 20;

    Staging()
    {
        for (int b = 0; b < 2; ++b)
            if (hipHostMalloc(&buf_[b], kChunk, hipHostMallocDefault) != hipSuccess) buf_[b] = nullptr;
    }
    ~Staging()
    {
        for (int b = 0; b < 2; ++b)
            if (buf_[b]) (void)hipHostFree(buf_[b]);
    }
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    bool ok() const { return buf_[0] && buf_[1]; }

    // bytes from host src to device dst on stream s; returns when they are
    // on the device (work queued on s before the call runs first)
    void upload(void* dst, const void* src, size_t bytes, hipStream_t s)
    {
        Events2 ev;
        bool used[2] = {false, false};
        for (size_t i = 0, off = 0; off < bytes; ++i, off += kChunk) {
            const int b = int(i & 1);
            const size_t len = std::min(kChunk, bytes - off);
            if (used[b]) check(hipEventSynchronize(ev.e[b]), "upload");  // its previous chunk has left
            parallel_copy(buf_[b], static_cast<const char*>(src) + off, len);
            check(hipMemcpyAsync(static_cast<char*>(dst) + off, buf_[b], len, hipMemcpyHostToDevice, s), "upload");
            check(hipEventRecord(ev.e[b], s), "upload");
            used[b] = true;
        }
        // the last chunks' copies must finish before their buffers are reused
        for (int b = 0; b < 2; ++b)
            if (used[b]) check(hipEventSynchronize(ev.e[b]), "upload");
    }

    // bytes from device src (ready once the work queued on s is done) to
    // host dst; returns when all of them are in dst
    void download(void* dst, const void* src, size_t bytes, hipStream_t s)
    {
        Events2 ev;
        const size_t chunks = (bytes + kChunk - 1) / kChunk;
        auto len_of = [&](size_t i) { return std::min(kChunk, bytes - i * kChunk); };
        auto issue = [&](size_t i) {
            const int b = int(i & 1);
            check(hipMemcpyAsync(buf_[b], static_cast<const char*>(src) + i * kChunk, len_of(i),
                                 hipMemcpyDeviceToHost, s),
                  "download");
            check(hipEventRecord(ev.e[b], s), "download");
        };
        if (chunks) issue(0);
        for (size_t i = 0; i < chunks; ++i) {
            // chunk i + 1 goes to the buffer whose chunk i - 1 was drained below
            if (i + 1 < chunks) issue(i + 1);
            const int b = int(i & 1);
            check(hipEventSynchronize(ev.e[b]), "download");
            parallel_copy(static_cast<char*>(dst) + i * kChunk, buf_[b], len_of(i));
        }
    }

private:
    struct Events2 {
        hipEvent_t e[2] = {nullptr, nullptr};
        Events2()
        {
            for (int b = 0; b < 2; ++b) {
                const hipError_t r = hipEventCreateWithFlags(&e[b], hipEventDisableTiming);
                if (r != hipSuccess) {
                    e[b] = nullptr;
                    if (b) (void)hipEventDestroy(e[0]);
                    check(r, "hipEventCreate");
                }
            }
        }
        ~Events2()
        {
            for (int b = 0; b < 2; ++b)
                if (e[b]) (void)hipEventDestroy(e[b]);
        }
    };
    void* buf_[2] = {nullptr, nullptr};
};

// The staging buffers of the host-pointer path, kept for the life of the
// process (one pair; a call that finds them in use by another host thread
// stages through a pair of its own).
std::mutex g_staging_mutex;
Staging* g_staging = nullptr;  // never freed: it lives until the process ends

// Runs fn(staging) with the shared staging buffers, or a private pair if
// another thread holds them, or with nullptr when pinned memory cannot be had.
template <class Fn>
void with_staging(Fn fn)
{
    std::unique_lock<std::mutex> lock(g_staging_mutex, std::try_to_lock);
    if (lock.owns_lock()) {
        if (!g_staging) {
            std::unique_ptr<Staging> st(new Staging());
            if (st->ok()) g_staging = st.release();
        }
        if (g_staging) {
            fn(g_staging);
            return;
        }
        lock.unlock();
    }
    Staging own;
    fn(own.ok() ? &own : nullptr);
}

void require_device()
{
    int devices = 0;
    const hipError_t e = hipGetDeviceCount(&devices);
    if (e != hipSuccess || devices == 0) throw ExecutionError("No HIP device is available");
}

// Records each lane's ring holds: every pair the minimiser keeps, which is
// never more than an inner solve's evaluations.
int history_cap(int memory) { return std::min(memory, kMaxEvals); }

// Blocks a history region holds for a launch over `chunks` pixel chunks of
// the default size (a launch never has more blocks than chunks; launches
// with smaller chunks are capped at this many blocks).
int history_blocks(long long chunks) { return int(std::max(1LL, std::min<long long>(kHistBlocks, chunks))); }

size_t history_bytes(long long chunks, int layers, int memory)
{
    if (memory == 1) return 0;
    return align256(size_t(history_blocks(chunks)) * 64 * size_t(history_cap(memory)) *
                    size_t(kHistRecord(layers)) * sizeof(double));
}

long long default_chunks(size_t n) { return (long long)((n + kDefaultChunk - 1) / kDefaultChunk); }

}  // namespace

void check_variants(const unblending_hip_params& p, int layers)
{
    if (p.sparsity != 0 && p.sparsity != 1) throw std::invalid_argument("sparsity must be 0 or 1");
    if (p.gray_layers == 0) return;
    if (p.gray_layers != 2)
        throw std::invalid_argument("The device solves the second-layer gray variant only (gray_layers = 2)");
    if (layers < 2) throw std::invalid_argument("A gray second layer needs at least 2 layers");
    if (lbfgs_memory(4 * layers, kMaxEvals, p.lbfgs_memory) == 1)
        throw std::invalid_argument("The gray variant runs with the default L-BFGS history (lbfgs_memory != 1)");
}

int lbfgs_memory(int n, int max_evals, int memory)
{
    if (memory > 0) return memory;
    int m = 1310720 / (n > 0 ? n : 1);
    if (m < 10) m = 10;
    if (max_evals > 0 && max_evals <= m) m = max_evals;
    return m;
}

bool read_params(const unblending_hip_params* params, unblending_hip_params& p)
{
    p = UNBLENDING_HIP_PARAMS_INIT;
    if (!params) return true;
    const size_t minimum = offsetof(unblending_hip_params, warm_outer_cap) + sizeof(int32_t);
    if (params->struct_size < minimum) return false;
    std::memcpy(&p, params, std::min<size_t>(params->struct_size, sizeof p));
    p.struct_size = sizeof p;
    return true;
}

// filter workspace: the bounds slots, then every channel's coefficient planes
constexpr size_t kBoundsBytes = 256;
static_assert(kBoundsBytes >= (1 + kMaxChannels) * sizeof(unsigned long long), "bounds slots");

size_t filter_workspace_bytes(size_t n, int channels)
{
    return kBoundsBytes + align256(4 * size_t(channels) * n * sizeof(double));
}

namespace {

// Carves `bytes` (256-aligned) from a workspace cursor.
template <class T>
T* carve(char*& cur, size_t count)
{
    T* p = reinterpret_cast<T*>(cur);
    cur += align256(count * sizeof(T));
    return p;
}
}  // namespace

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
    b.partial = carve<float>(cur, 4 * size_t(mlp::kGradSplits) * kHidden * kHidden);
    b.loss = carve<float>(cur, mlp::kLossSlots);
    return b;
}

size_t mlp_train_bytes(int batch)
{
    // a dry run of the carving over a null base gives the size
    const mlp::TrainBuffers b = mlp_train_view(nullptr, batch);
    return size_t(reinterpret_cast<char*>(b.loss) - static_cast<char*>(nullptr)) + align256(mlp::kLossSlots * sizeof(float));
}

// Workspace of one pass: counters and descriptor, the per-lane outer
// iterates, (refinement) the filter's coefficient and output planes, then
// the (s, y) history.  The warm start's prediction needs none.
size_t workspace_bytes(size_t n, size_t layers, bool refine, bool smooth, const unblending_hip_params& params)
{
    size_t bytes = kHeader + align256(size_t(kLaneStateDoubles) * sizeof(double));
    if (refine) {
        const int C = filter_channels(layers, smooth);
        bytes += filter_workspace_bytes(n, C);
        bytes += align256(size_t(C) * n * sizeof(double));
    }
    const int M = lbfgs_memory(int(4 * layers), kMaxEvals, params.lbfgs_memory);
    return bytes + history_bytes(default_chunks(n), int(layers), M);
}

void run_filter(const double* rgb, int w, int h, const Channel* in, double* const* q, int channels, int radius,
                double eps, double* ab, hipStream_t stream, unblending_hip_stats* stats, bool timing, int block,
                int rows)
{
    if (radius > kMaxFilterRadius)
        throw ExecutionError("Guided-filter radius is outside the device filter's domain");
    if (channels > kMaxChannels) throw ExecutionError("too many filtered channels");
    Events ea(timing), eb(timing);
    double coeff_ms = 0.0, out_ms = 0.0;
    int launches = 0;
    // ab points at the filter workspace: bounds slots, then coefficient planes
    auto* bounds = reinterpret_cast<unsigned long long*>(ab);
    double* planes = reinterpret_cast<double*>(reinterpret_cast<char*>(ab) + kBoundsBytes);
    if (radius > 0) {
        BoundsArgs ba{};
        ba.rgb = rgb;
        ba.w = w;
        ba.h = h;
        ba.channels = channels;
        for (int c = 0; c < channels; ++c) ba.in[c] = in[c];
        ba.bounds = bounds;
        ea.start(stream);
        check(launch_zero(bounds, 1 + kMaxChannels, stream), "guided filter (bounds)");
        check(launch_filter_bounds(ba, stream), "guided filter (bounds)");
        ea.stop(stream);
        launches += 2;
        if (timing) coeff_ms += ea.ms();
    }
    FilterArgs fa{};
    fa.rgb = rgb;
    fa.w = w;
    fa.h = h;
    fa.r = radius;
    fa.eps = eps;
    fa.isr = 1.0 / std::sqrt(eps);
    fa.channels = channels;
    for (int c = 0; c < channels; ++c) {
        fa.in[c] = in[c];
        fa.q[c] = q[c];
    }
    fa.ab = planes;
    fa.bounds = bounds;
    fa.block = block;
    fa.rows = rows;
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
    args.sp = default_solver(job.params, int(L), refine);
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
        const mlp::Model md = mlp::model_view(const_cast<void*>(job.params.mlp));
        Events ep(timing);
        ep.start(job.stream);
        // the prediction is written into the output buffer, which the solve
        // reads pixel by pixel before writing the same pixel's solution
        check(mlp_predict(md, job.rgb, job.w, job.h, int(4 * L), job.out, job.stream), "warm-start prediction");
        ep.stop(job.stream);
        predict_ms = ep.ms();
        args.warm = job.out;
    }
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
        run_filter(job.rgb, job.w, job.h, in.data(), q.data(), C, radius, eps, ab, job.stream, job.stats, timing,
                   job.params.filter_block, job.params.filter_rows);
        args.first = job.first;
        args.filtered = qbase;
    }

    if (warm && args.solver != 0) throw std::invalid_argument("The warm start runs on the persistent solve kernel");
    if (args.sp.memory != 1 && args.solver != 0)
        throw std::invalid_argument("The one-pixel-per-lane kernel keeps one (s, y) pair (lbfgs_memory = 1)");
    check_variants(job.params, int(L));
    if (args.solver != 0 && (job.params.gray_layers || job.params.sparsity))
        throw std::invalid_argument("The problem variants run on the persistent solve kernel");
    ImageDesc desc{};
    BatchArgs batch{};
    ImageDesc* ddesc = reinterpret_cast<ImageDesc*>(ws + kDescOffset);
    if (args.solver == 0) {
        for (size_t l = 0; l < L; ++l) desc.layers[l] = args.layers[l];
        desc.rgb = args.rgb;
        desc.first = args.first;
        desc.filtered = args.filtered;
        desc.out = args.out;
        desc.alpha_planes = args.alpha_planes;
        desc.warm = args.warm;
        desc.fallback_count = counters + kFallback;
        desc.n = args.n;
        desc.L = int(L);
        desc.opaque = args.opaque;
        desc.smooth = args.smooth;
        desc.all_plus = args.all_plus;
        desc.all_so = args.all_so;
        check(launch_store_desc(ddesc, desc, job.stream), "image descriptor");
        batch.images = ddesc;
        batch.image_count = 1;
        batch.chunk_size = job.params.chunk_size > 0 ? job.params.chunk_size : kDefaultChunk;
        batch.chunk_count = (args.n + batch.chunk_size - 1) / batch.chunk_size;
        batch.sp = args.sp;
        batch.lane_state = args.lane_state;
        batch.counters = counters;
        batch.queue = kQueue;
        if (args.sp.memory != 1) {
            const size_t hbytes = history_bytes(default_chunks(n), int(L), args.sp.memory);
            batch.hist = reinterpret_cast<double*>(ws + workspace_bytes(n, L, refine, job.smooth, job.params) - hbytes);
            batch.hist_cap = history_cap(args.sp.memory);
            batch.hist_blocks = history_blocks(default_chunks(n));
        }
    }

    Events es(timing);
    es.start(job.stream);
    if (args.solver == 0) check(launch_batch_solve(int(L), refine, batch, job.stream), "solve");
    else check(launch_solve(int(L), refine, args, job.stream), "solve");
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
        job.stats->lane_slots = host_counters[kSlots];
        job.stats->history_pairs = host_counters[kPairs];
        job.stats->solve_ms = es.ms();
        job.stats->predict_ms = predict_ms;
    }
}

namespace {

// Workspace of a batch call: counters, one descriptor per image, the solve's
// lane state, the filter's coefficient planes (shared by the images, filtered
// one after another) and every image's filtered planes (image i's start
// q_offset[i] doubles into the planes area).
struct BatchLayout {
    size_t desc, lane_state, ab, planes, hist, total;
    size_t ab_stride = 0;  // one coefficient region per concurrent filter stream
    int filter_streams = 1;
    std::vector<size_t> q_offset;
    int hist_blocks = 0, hist_cap = 0;
};

// Images of a batch filtered at once, each stream with its own coefficient
// planes: a batch of small images (c5: 64 images of 0.03-4 Mpixel) fills only
// part of the device with one image's filter launches.
constexpr int kFilterStreams = 4;

BatchLayout batch_layout(const unblending_hip_image* images, size_t count, bool refine, bool smooth,
                         const unblending_hip_params& params)
{
    BatchLayout b;
    b.desc = kDescOffset;
    b.lane_state = align256(b.desc + count * kDescStride);
    b.ab = b.lane_state + align256(size_t(kLaneStateDoubles) * sizeof(double));
    size_t ab_bytes = 0, planes = 0;
    b.q_offset.assign(count, 0);
    if (refine) {
        for (size_t i = 0; i < count; ++i) {
            const size_t n = size_t(images[i].width) * size_t(images[i].height);
            const int C = filter_channels(images[i].layer_count, smooth);
            ab_bytes = std::max(ab_bytes, filter_workspace_bytes(n, C));
            b.q_offset[i] = planes;
            planes += (size_t(C) * n + 31) / 32 * 32;
        }
    }
    b.filter_streams = refine ? int(std::min<size_t>(kFilterStreams, std::max<size_t>(count, 1))) : 1;
    b.ab_stride = align256(ab_bytes);
    b.planes = b.ab + size_t(b.filter_streams) * b.ab_stride;
    b.hist = b.planes + align256(planes * sizeof(double));
    // one history region, shared by the launches of the batch (one after
    // another on the stream), sized for the largest layer count's records
    long long chunks = 0;
    size_t hist = 0;
    for (size_t i = 0; i < count; ++i) chunks += default_chunks(size_t(images[i].width) * size_t(images[i].height));
    for (size_t i = 0; i < count; ++i) {
        const int L = int(images[i].layer_count);
        const int M = lbfgs_memory(4 * L, kMaxEvals, params.lbfgs_memory);
        hist = std::max(hist, history_bytes(chunks, L, M));
        if (M != 1) b.hist_cap = std::max(b.hist_cap, history_cap(M));
    }
    b.hist_blocks = history_blocks(chunks);
    b.total = b.hist + hist;
    return b;
}

}  // namespace

// The streams a batch's filters run on besides the caller's, and the events
// that fork them from and join them to the caller's stream: one set per
// device, created at the first batch refinement on it and kept until the
// process ends (never created during a graph capture: a batch captured before
// its device has them filters on the caller's stream).  The mutex keeps one
// batch at a time between a fork and its join.
constexpr int kMaxDevices = 64;
struct FilterSide {
    std::mutex mutex;
    hipStream_t s[kFilterStreams] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kFilterStreams] = {};
};

FilterSide* filter_side(hipStream_t caller)
{
    static std::mutex sides_mutex;
    static FilterSide* sides[kMaxDevices] = {};  // never freed
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> guard(sides_mutex);
    if (sides[dev]) return sides[dev];
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(caller, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
    std::unique_ptr<FilterSide> f(new FilterSide());
    bool ok = hipEventCreateWithFlags(&f->fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < kFilterStreams; ++k)
        ok = hipStreamCreateWithFlags(&f->s[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&f->join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        for (int k = 0; k < kFilterStreams; ++k) {
            if (f->s[k]) (void)hipStreamDestroy(f->s[k]);
            if (f->join[k]) (void)hipEventDestroy(f->join[k]);
        }
        if (f->fork) (void)hipEventDestroy(f->fork);
        return nullptr;
    }
    sides[dev] = f.release();
    return sides[dev];
}

size_t batch_workspace_bytes(const unblending_hip_image* images, size_t count, bool refine, bool smooth,
                             const unblending_hip_params& params)
{
    return batch_layout(images, count, refine, smooth, params).total;
}

void run_batch(const unblending_hip_image* images, const std::vector<LayerSet>& sets, bool refine, bool opaque,
               bool smooth, const unblending_hip_params& params, void* workspace, hipStream_t stream,
               unblending_hip_stats* stats)
{
    const size_t count = sets.size();
    const bool timing = params.timing != 0;
    const BatchLayout lay = batch_layout(images, count, refine, smooth, params);
    char* ws = static_cast<char*>(workspace);
    auto* counters = reinterpret_cast<unsigned long long*>(ws);
    ImageDesc* ddesc = reinterpret_cast<ImageDesc*>(ws + lay.desc);
    double* planes = reinterpret_cast<double*>(ws + lay.planes);
    if (stats) *stats = unblending_hip_stats{};
    check(launch_zero(counters, kNumCounters, stream), "zero counters");

    // first pass with a warm start: each image with a model gets that model's
    // prediction in its output buffer, which its solve reads pixel by pixel
    // before writing the same pixel's solution
    const bool warm = !refine && params.warm_start != 0;
    if (warm) {
        Events ep(timing);
        ep.start(stream);
        for (size_t i = 0; i < count; ++i) {
            const unblending_hip_image& im = images[i];
            if (!im.mlp) continue;
            const mlp::Model md = mlp::model_view(const_cast<void*>(im.mlp));
            check(mlp_predict(md, im.rgb, im.width, im.height, int(4 * sets[i].host.size()), im.output, stream),
                  "warm-start prediction");
        }
        ep.stop(stream);
        if (stats) stats->predict_ms = ep.ms();
    }

    // refinement: filter every image's alphas (and background) into its
    // planes, up to kFilterStreams images at once: each image goes to the
    // stream with the fewest pixels so far (largest images first), with that
    // stream's coefficient region
    if (refine) {
        FilterSide* side = lay.filter_streams > 1 ? filter_side(stream) : nullptr;
        std::unique_lock<std::mutex> side_lock;
        if (side) side_lock = std::unique_lock<std::mutex>(side->mutex);
        const int lanes = side ? lay.filter_streams : 1;
        std::vector<size_t> by_size(count);
        for (size_t i = 0; i < count; ++i) by_size[i] = i;
        auto pixels = [&](size_t i) { return size_t(images[i].width) * size_t(images[i].height); };
        std::stable_sort(by_size.begin(), by_size.end(), [&](size_t a, size_t b) { return pixels(a) > pixels(b); });
        std::vector<size_t> load(lanes, 0);
        Events ew(timing);
        ew.start(stream);
        if (side) {
            check(hipEventRecord(side->fork, stream), "filter fork");
            for (int k = 0; k < lanes; ++k) check(hipStreamWaitEvent(side->s[k], side->fork, 0), "filter fork");
        }
        unblending_hip_stats fs{};
        int launches = 0;
        for (size_t i : by_size) {
            const int k = int(std::min_element(load.begin(), load.end()) - load.begin());
            load[k] += pixels(i);
            hipStream_t fstream = side ? side->s[k] : stream;
            double* ab = reinterpret_cast<double*>(ws + lay.ab + size_t(k) * lay.ab_stride);
            const unblending_hip_image& im = images[i];
            const size_t n = pixels(i), L = sets[i].host.size();
            int radius = im.filter_radius;
            if (radius < 0) radius = 60 * std::min(im.width, im.height) / 1000;  // unblending.cpp:351
            const double eps = params.filter_epsilon > 0.0 ? params.filter_epsilon : 1e-4;  // :353
            const int C = filter_channels(L, smooth);
            std::vector<Channel> in(C);
            std::vector<double*> q(C);
            for (size_t l = 0; l < L; ++l)
                in[l] = im.alpha_planes ? Channel{im.alpha_planes + l * n, 1} : Channel{im.first_pass + 4 * l * n + 3, 4};
            for (int c = 0; c < (smooth ? 3 : 0); ++c) in[L + c] = Channel{im.first_pass + c, 4};
            for (int c = 0; c < C; ++c) q[c] = planes + lay.q_offset[i] + size_t(c) * n;
            // untimed per image (a per-image timing would wait for each
            // filter in turn): the phase is timed as a whole below
            run_filter(im.rgb, im.width, im.height, in.data(), q.data(), C, radius, eps, ab, fstream,
                       stats ? &fs : nullptr, false, params.filter_block, params.filter_rows);
            launches += fs.filter_launches;
        }
        if (side) {
            for (int k = 0; k < lanes; ++k) {
                check(hipEventRecord(side->join[k], side->s[k]), "filter join");
                check(hipStreamWaitEvent(stream, side->join[k], 0), "filter join");
            }
        }
        ew.stop(stream);
        if (stats) {
            stats->filter_ms = timing ? ew.ms() : 0.0;
            stats->filter_coeff_ms = 0.0;  // the passes of concurrent images are not timed apart
            stats->filter_out_ms = 0.0;
            stats->filter_launches = launches;
        }
    }

    // descriptors, grouped by layer count (stable)
    std::vector<size_t> order(count);
    for (size_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return sets[a].host.size() < sets[b].host.size(); });
    const int chunk = params.chunk_size > 0 ? params.chunk_size : kDefaultChunk;
    for (size_t i = 0; i < count; ++i) check_variants(params, int(sets[i].host.size()));
    struct Group {
        int L;
        size_t first, count;
        long long chunks;
    };
    std::vector<Group> groups;
    long long chunk_base = 0;
    for (size_t k = 0; k < count; ++k) {
        const size_t i = order[k];
        const unblending_hip_image& im = images[i];
        const LayerSet& set = sets[i];
        const int L = int(set.host.size());
        if (groups.empty() || groups.back().L != L) {
            groups.push_back(Group{L, k, 0, 0});
            chunk_base = 0;
        }
        ImageDesc d{};
        for (int l = 0; l < L; ++l) d.layers[l] = to_device(set.host[l]);
        d.rgb = im.rgb;
        d.n = (long long)im.width * im.height;
        d.out = im.output;
        if (refine) {
            d.first = im.first_pass;
            d.filtered = planes + lay.q_offset[i];
        } else {
            d.alpha_planes = im.alpha_planes;
            if (warm && im.mlp) d.warm = im.output;
            d.fallback_count = counters + kFallback;
        }
        d.chunk_base = chunk_base;
        d.L = L;
        d.opaque = opaque ? 1 : 0;
        d.smooth = (refine && smooth) ? 1 : 0;
        d.all_plus = set.all_plus ? 1 : 0;
        d.all_so = set.all_so ? 1 : 0;
        check(launch_store_desc(ddesc + k, d, stream), "image descriptor");
        const long long chunks = (d.n + chunk - 1) / chunk;
        chunk_base += chunks;
        groups.back().count += 1;
        groups.back().chunks += chunks;
    }
    if (groups.size() > size_t(kNumCounters - kGroupQueue))
        throw ExecutionError("too many distinct layer counts in one batch");

    Events es(timing);
    es.start(stream);
    for (size_t g = 0; g < groups.size(); ++g) {
        BatchArgs a{};
        a.images = ddesc + groups[g].first;
        a.image_count = int(groups[g].count);
        a.chunk_size = chunk;
        a.chunk_count = groups[g].chunks;
        a.sp = default_solver(params, groups[g].L, refine);
        a.lane_state = reinterpret_cast<double*>(ws + lay.lane_state);
        a.counters = counters;
        a.queue = kGroupQueue + int(g);
        if (a.sp.memory != 1) {
            a.hist = reinterpret_cast<double*>(ws + lay.hist);
            a.hist_cap = history_cap(a.sp.memory);
            a.hist_blocks = lay.hist_blocks;
        }
        check(launch_batch_solve(groups[g].L, refine, a, stream), "batch solve");
    }
    es.stop(stream);
    if (stats) {
        unsigned long long host_counters[kNumCounters] = {};
        check(hipMemcpyAsync(host_counters, counters, sizeof host_counters, hipMemcpyDeviceToHost, stream),
              "counters");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        stats->evals = host_counters[kEvals];
        stats->iterations = host_counters[kIterations];
        stats->outer = host_counters[kOuter];
        stats->pixels = host_counters[kPixels];
        stats->lane_slots = host_counters[kSlots];
        stats->history_pairs = host_counters[kPairs];
        stats->fallback_pixels = host_counters[kFallback];
        stats->solve_ms = es.ms();
    }
}

namespace {

// Device buffers of the host-pointer path (the solve's image, result, first
// pass and workspace; the recomposite's layers and result), kept between
// calls on each device and grown when a call needs more.  Allocated and freed at every call they cost the driver's clearing
// of the memory at each allocation once its cleared memory is used up:
// measured on an MI355X, a 1024^2, 5-layer unmix went from 70 ms to 620 ms
// a call after its first calls (its workspace holds a 22 GB (s, y) history
// region).  A call that finds them in use by another host thread allocates
// its own; unblending_hip_release_buffers frees them.
constexpr int kHostBufferKinds = 6;  // solve: image, result, first pass, workspace; composite: layers, result
struct HostBuffers {
    std::mutex mutex;
    void* p[kHostBufferKinds] = {};
    size_t bytes[kHostBufferKinds] = {};
};
std::mutex g_host_buffers_mutex;
HostBuffers* g_host_buffers[kMaxDevices] = {};  // never freed: they live until the process ends

HostBuffers* host_buffers()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> guard(g_host_buffers_mutex);
    if (!g_host_buffers[dev]) g_host_buffers[dev] = new HostBuffers();
    return g_host_buffers[dev];
}

// One device buffer of a host-pointer call: the device's kept one (grown to
// `bytes` when smaller) if the call holds their lock, else one of its own.
class HostBuffer {
public:
    HostBuffer(HostBuffers* hb, bool held, int kind, size_t bytes)
    {
        if (!bytes) return;
        if (!held) {
            own_.reset(new DeviceBuffer(bytes));
            p_ = own_->p;
            return;
        }
        if (hb->bytes[kind] < bytes) {
            if (hb->p[kind]) (void)hipFree(hb->p[kind]);
            hb->p[kind] = nullptr;
            hb->bytes[kind] = 0;
            const hipError_t e = hipMalloc(&hb->p[kind], bytes);
            if (e != hipSuccess) {
                hb->p[kind] = nullptr;
                if (e == hipErrorOutOfMemory) throw std::bad_alloc();
                check(e, "hipMalloc");
            }
            hb->bytes[kind] = bytes;
        }
        p_ = hb->p[kind];
    }
    void* get() const { return p_; }
    template <class T> T* as() const { return static_cast<T*>(p_); }

private:
    void* p_ = nullptr;
    std::unique_ptr<DeviceBuffer> own_;
};

}  // namespace

void release_host_buffers()
{
    HostBuffers* hb = host_buffers();
    if (!hb) return;
    std::lock_guard<std::mutex> guard(hb->mutex);  // after a call in progress
    for (int k = 0; k < kHostBufferKinds; ++k) {
        if (hb->p[k]) (void)hipFree(hb->p[k]);
        hb->p[k] = nullptr;
        hb->bytes[k] = 0;
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
    HostBuffers* hb = host_buffers();
    std::unique_lock<std::mutex> held;
    if (hb) held = std::unique_lock<std::mutex>(hb->mutex, std::try_to_lock);
    const bool keep = held.owns_lock();
    HostBuffer d_rgb(hb, keep, 0, 3 * n * sizeof(double));
    HostBuffer d_out(hb, keep, 1, 4 * L * n * sizeof(double));
    HostBuffer d_first(hb, keep, 2, first ? 4 * L * n * sizeof(double) : 0);
    const size_t ws_bytes = workspace_bytes(n, L, first != nullptr, smooth, params);
    HostBuffer d_ws(hb, keep, 3, ws_bytes);
    check(hipMemcpy2DAsync(d_rgb.get(), 3 * w * sizeof(double), rgb.data, rgb.row_stride * sizeof(double),
                           3 * w * sizeof(double), h, hipMemcpyHostToDevice, stream.s),
          "upload rgb");
    if (first) {
        with_staging([&](Staging* st) {
            if (st) st->upload(d_first.get(), first, 4 * L * n * sizeof(double), stream.s);
            else
                check(hipMemcpyAsync(d_first.get(), first, 4 * L * n * sizeof(double), hipMemcpyHostToDevice, stream.s),
                      "upload first pass");
        });
    }
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
    job.workspace = d_ws.get();
    job.workspace_size = ws_bytes;
    job.stream = stream.s;
    job.stats = stats;
    run_device(job);
    check(hipStreamSynchronize(stream.s), "solve");
    with_staging([&](Staging* st) {
        if (st) {
            st->download(output, d_out.get(), 4 * L * n * sizeof(double), stream.s);
        } else {
            check(hipMemcpyAsync(output, d_out.get(), 4 * L * n * sizeof(double), hipMemcpyDeviceToHost, stream.s),
                  "download");
            check(hipStreamSynchronize(stream.s), "download");
        }
    });
}

// Image-level recomposite (unblending_hip_composite*): comp_ops [count][3]
// and modes [count] are validated by the caller.
void run_composite(const double* layers, int w, int h, size_t count, const int32_t* comp_ops, const int32_t* modes,
                   double* out, hipStream_t stream)
{
    CompositeArgs a{};
    for (size_t k = 0; k < count; ++k) {
        a.table[k].X = double(comp_ops[3 * k]);
        a.table[k].Y = double(comp_ops[3 * k + 1]);
        a.table[k].Z = double(comp_ops[3 * k + 2]);
        a.table[k].mode = modes[k];
    }
    a.count = int(count);
    a.layers = layers;
    a.out = out;
    a.n = (long long)w * h;
    check(launch_composite(a, stream), "composite");
}

void host_composite(const double* layers, int w, int h, size_t count, const int32_t* comp_ops, const int32_t* modes,
                    double* out)
{
    require_device();
    const size_t n = size_t(w) * size_t(h);
    Stream stream;
    HostBuffers* hb = host_buffers();
    std::unique_lock<std::mutex> held;
    if (hb) held = std::unique_lock<std::mutex>(hb->mutex, std::try_to_lock);
    const bool keep = held.owns_lock();
    HostBuffer d_in(hb, keep, 4, 4 * count * n * sizeof(double));
    HostBuffer d_out(hb, keep, 5, 4 * n * sizeof(double));
    with_staging([&](Staging* st) {
        if (st) st->upload(d_in.get(), layers, 4 * count * n * sizeof(double), stream.s);
        else
            check(hipMemcpyAsync(d_in.get(), layers, 4 * count * n * sizeof(double), hipMemcpyHostToDevice, stream.s),
                  "upload layers");
    });
    run_composite(d_in.as<double>(), w, h, count, comp_ops, modes, d_out.as<double>(), stream.s);
    with_staging([&](Staging* st) {
        if (st) {
            st->download(out, d_out.get(), 4 * n * sizeof(double), stream.s);
        } else {
            check(hipMemcpyAsync(out, d_out.get(), 4 * n * sizeof(double), hipMemcpyDeviceToHost, stream.s), "download");
            check(hipStreamSynchronize(stream.s), "composite");
        }
    });
}

}  // namespace ubhip
