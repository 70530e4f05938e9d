// ub_kernels.hpp -- launch interface between the host pipeline and the gfx950
// kernels (types shared by ub_solve.hip, ub_filter.hip and ub_pipeline.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "ub_device.hpp"

namespace ubhip {

constexpr int kMaxLayers = 8;      // layer counts with a compiled solve kernel
constexpr int kLbfgsMemory = 1;    // (s, y) pairs of the register-resident inner minimiser
constexpr int kMaxPersistentBlocks = 2048;  // 64-lane blocks of the persistent solve
constexpr long long kLaneStateDoubles = (long long)kMaxPersistentBlocks * 64 * 4 * kMaxLayers;
// 64-lane blocks of a persistent solve that keeps its (s, y) history in the
// workspace: one wave per SIMD on the 256 CUs of an MI355X.  Each lane owns
// a ring of min(memory, max_evals) records of kHistRecord(L) doubles: s and
// y (4L components each), then 1 / s.y and the two-loop multiplier a.  A
// record is a multiple of 16 bytes, so s and y start 16-byte aligned.
constexpr int kHistBlocks = 1024;
constexpr int kHistRecord(int L) { return 8 * L + 2; }

// Solver constants, unblending/src/unblending.cpp:220-249, plus the inner
// minimiser's own (line-search trials).
struct SolverParams {
    int max_evals;   // 1000
    int max_count;   // 20: the outer loop stops when count > max_count
    int max_ls;      // 20
    int alpha_fixed; // refinement alphas fixed by bounds instead of constraints
    int warm_max_count;  // warm-started attempt: outer cap before the pixel is solved from the default start
    int memory;      // (s, y) pairs kept: 1 = the register-resident minimiser, else a history in the workspace
    int gray;        // 1: layer 1's colour constrained to gray (history-keeping kernel, 2 to 8 layers)
    double ftol, xtol, eps, gamma, beta, rho0;
    double sparse_sigma;  // the sparsity term's weight in the energy, 0 = off (first pass only)
};

// Per-pass work counters accumulated by the solve kernels.
// kQueue / kQueue2 are the persistent solve's next-item counters; kFallback
// counts the pixels whose warm-started attempt missed the tolerances and
// that were solved again from the default start.
// kSlots counts the persistent solve's lane evaluations, useful or idle
// (trips of its loop times 64).
enum Counter : int {
    kEvals = 0, kIterations = 1, kOuter = 2, kPixels = 3, kQueue = 4, kFallback = 5, kQueue2 = 6, kSlots = 7,
    kPairs = 8,        // (s, y) pairs the two-loop recursions went through
    kGroupQueue = 16,  // batches: the queue of the launch of the g-th layer count is kGroupQueue + g
    kNumCounters = 24
};

struct PixelArgs {
    DevLayer layers[kMaxLayers];  // passed by value; read by the kernel with scalar loads
    const double* rgb;     // [n][3] target colours
    long long n;           // pixels
    SolverParams sp;
    int opaque, smooth, all_plus, all_so;
    int solver;            // 0 = persistent kernel, 1 = one pixel per lane
    const double* first;   // refinement: first pass, packed [L][n][RGBA]
    const double* filtered;// refinement: guided-filter output planes [L (+3)][n]
    double* out;           // packed [L][n][RGBA]
    double* alpha_planes;  // optional copy of the solved alphas [L][n]
    double* lane_state;    // persistent solve: kLaneStateDoubles of per-lane scratch
    unsigned long long* counters;
    // first pass only: warm start
    const double* warm;    // predicted first pass, packed [L][n][RGBA], or nullptr
};

// One-pixel-per-lane solve of one image (PixelArgs::solver = 1).
hipError_t launch_solve(int L, bool refine, const PixelArgs& args, hipStream_t stream);

// One image of a persistent solve launch: its layer set, buffers and the
// range of the launch's pixel chunks it owns.  Lives in device memory and is
// read-only during the launch.
struct ImageDesc {
    DevLayer layers[kMaxLayers];
    const double* rgb;       // [n][3]
    const double* first;     // refinement: first pass, packed [L][n][RGBA]
    const double* filtered;  // refinement: guided-filter output planes [L (+3)][n]
    double* out;             // packed [L][n][RGBA]
    double* alpha_planes;    // optional copy of the solved alphas [L][n]
    const double* warm;      // first pass: predicted first pass (warm start) or nullptr
    unsigned long long* fallback_count;  // warm start: pixels solved again from the default start
    const long long* list;   // solve list[0 .. *list_count) instead of pixels 0 .. n-1
    const unsigned long long* list_count;
    long long n;             // pixels
    long long chunk_base;    // the image's first chunk; its chunks are consecutive
    int L, opaque, smooth, all_plus, all_so, pad;
};

// A persistent solve over a batch of images (one pass kind).  The images'
// pixels are cut into chunks of chunk_size (chunk k of image i covers pixels
// (k - chunk_base) * chunk_size ...); a wave takes chunks from the counter
// counters[queue] and solves one image at a time.
struct BatchArgs {
    const ImageDesc* images;
    int image_count;
    int chunk_size;
    long long chunk_count;
    SolverParams sp;
    double* lane_state;      // kLaneStateDoubles of per-lane scratch
    unsigned long long* counters;
    int queue;
    int max_blocks;          // 0: as many as fit on the device
    // sp.memory != 1: the (s, y) history, hist_blocks blocks of 64 lanes x
    // hist_cap records of kHistRecord(L) doubles, [block][lane][record]
    double* hist;
    int hist_cap;
    int hist_blocks;
};

constexpr int kDefaultChunk = 64;  // pixels a wave takes from the queue at once

// Every image of the batch must have L layers.
hipError_t launch_batch_solve(int L, bool refine, const BatchArgs& args, hipStream_t stream);
// The same for the problem variants (args.sp.gray, args.sp.sparse_sigma):
// history-keeping kernels only.
hipError_t launch_batch_variant(int L, bool refine, const BatchArgs& args, hipStream_t stream);

// One filtered channel: value(pixel) = base[pixel * stride].
struct Channel {
    const double* base;
    int stride;
};

constexpr int kMaxGroup = 3;      // channels of one launch of the wide strip kernel (r > 96)
constexpr int kMaxChannels = 16;  // channels of one filter call (L alphas + 3 background)
constexpr int kWaveStripMaxRadius = 96;  // radii filtered by the one-wave strip kernel

struct FilterArgs {
    const double* rgb;    // guidance [n][3]
    int w, h, r;
    double eps;
    double isr;           // 1 / sqrt(eps), the coefficient bounds' factor (computed on the host)
    int channels;         // channels of the call
    int first_channel;    // wide strip kernel: this launch's first channel
    int groups;           // wide strip kernel: channels in this launch (<= kMaxGroup)
    Channel in[kMaxChannels];  // pass A inputs
    double* ab;           // pass A output / pass B input: [4 * channels][n] (a_r, a_g, a_b, b per channel)
    double* q[kMaxChannels];   // pass B outputs
    const unsigned long long* bounds;  // max|I| then max|p_c| of the call's channels (bit patterns)
    int block;            // r <= 96: strip width in columns (128 or 256; 0: 256);
                          // r > 96: workgroup width (256, 512 or 1024; 0: chosen by radius)
    int rows;             // rows per strip (0: chosen by image size)
};

// max|I| and max|p_c| of every channel of a filter call: bounds[0] and
// bounds[1 + c] (bit patterns of non-negative doubles, zeroed beforehand).
struct BoundsArgs {
    const double* rgb;
    int w, h;
    int channels;
    Channel in[kMaxChannels];
    unsigned long long* bounds;
};

// The exact-sum filter (ub_filter.hip): the bounds pass, then pass A (box
// statistics and per-pixel coefficients of every channel) and pass B (box of
// the coefficients and the output q of every channel).  The passes return
// hipErrorInvalidValue when the radius is beyond their reach
// (r > kMaxFilterRadius).
constexpr int kMaxFilterRadius = 480;
hipError_t launch_filter_bounds(const BoundsArgs& args, hipStream_t stream);
hipError_t launch_filter_coefficients(const FilterArgs& args, hipStream_t stream);
hipError_t launch_filter_output(const FilterArgs& args, hipStream_t stream);

// Image-level recomposite (ub_composite.hip): packed [count][n][RGBA] layers
// folded bottom to top into [n][RGBA].
constexpr int kMaxCompositeLayers = 16;
struct CompositeLayer {
    double X, Y, Z;
    int mode;
    int pad;
};
struct CompositeArgs {
    CompositeLayer table[kMaxCompositeLayers];
    int count;
    const double* layers;
    double* out;
    long long n;
};
hipError_t launch_composite(const CompositeArgs& args, hipStream_t stream);

}  // namespace ubhip
