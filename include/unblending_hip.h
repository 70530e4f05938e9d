/*
 * unblending_hip.h -- extended entry points of the HIP backend.
 *
 * These have no one-to-one counterpart in the reference's public C ABI; each
 * one states the reference interface it generalises.
 *
 * Host-pointer variants (same buffers and validation as unblending.h):
 *   unblending_hip_unmix_ex / unblending_hip_refine_ex generalise
 *   unblending_unmix / unblending_refine (unblending/sdk/unblending_c.cpp:50-110)
 *   with the filter radius and epsilon that the reference hard-codes
 *   (unblending/src/unblending.cpp:351-353), the reference's
 *   MATTE_ALPHA_FORCE_FIXED variant (unblending.cpp:177-183) and work counters.
 *
 * Device-resident variants take device pointers on the current HIP device
 * and a caller workspace, enqueue on `stream` (a hipStream_t; NULL = the null
 * stream) and allocate nothing.  They are the seam the reference's
 * gpu::solve_cuda occupies (unblending/src/cuda/pixel_pipeline.cu:33-107,
 * declared at unblending/src/gpu_dispatch.hpp:43-45), with packed instead of
 * planar buffers:
 *   rgb         [height][width][3]   (tightly packed)
 *   first/out   [layer][height][width][RGBA]
 *   alpha_planes[layer][height][width] (optional by-product / input)
 *
 * uhip_solve_planes is that seam with the reference's planar layout
 * (gpu_dispatch.cpp:136-158: rgb[3][n], first[4L][n] alpha planes then colour
 * planes, out[4L][n]), host pointers.
 */
#ifndef UNBLENDING_HIP_H
#define UNBLENDING_HIP_H

#include "unblending.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct unblending_hip_params {
    uint32_t struct_size;
    int32_t filter_radius;    /* < 0: 60 * min(width, height) / 1000 */
    double filter_epsilon;    /* <= 0: 1e-4 */
    int32_t alpha_fixed;      /* refinement: 1 = alphas fixed at the filtered targets by bounds */
    int32_t solver;           /* 0 = default kernel, 1 = simple one-pixel-per-lane kernel */
    int32_t timing;           /* 1 = record per-kernel device time in the stats (synchronises) */
    int32_t warm_start;       /* first pass: 1 = start from the model's prediction (mlp); a pixel whose
                                 warm-started solve has not met the tolerances after warm_outer_cap + 1
                                 outer iterations is solved again from the default initialisation */
    const void* mlp;          /* device buffer of a trained model (unblending_hip_mlp_*) */
    int32_t chunk_size;       /* pixels a wave takes from the solve queue at once; 0: default */
    int32_t filter_block;     /* guided filter strip width: radius <= 96 one-wave strips of 128 or 256
                                 columns, larger radii workgroups of 256, 512 or 1024 threads;
                                 0: chosen by radius */
    int32_t filter_rows;      /* guided filter rows per strip; 0: chosen by image size */
    int32_t warm_outer_cap;   /* see warm_start; 0: 8 */
    int32_t lbfgs_memory;     /* (s, y) pairs the inner minimiser keeps.  0: NLopt's default vector
                                 storage for the reference's LD_LBFGS (unblending.cpp:249), which at
                                 maxeval 1000 keeps every pair of an inner solve; the history lives in
                                 the workspace.  1: one pair in registers, the fast mode (its results
                                 differ from the default's).  n > 1: n pairs. */
    int32_t gray_layers;      /* bit l: layer l's colour constrained to gray (sqrt(3) c - |c| 1 = 0), the
                                 reference's SECOND_LAYER_GRAY build (unblending.cpp:21,168-172), which
                                 is bit 1; the device solves 0 and bit 1 (2 to 8 layers, lbfgs_memory
                                 != 1).  0: none. */
    int32_t sparsity;         /* 1: the sparsity term 10 (sum alpha / sum alpha^2 - 1) in the first
                                 pass's energy, the reference's SPARCITY build (unblending.cpp:17,
                                 234-240).  0: off. */
} unblending_hip_params;

#define UNBLENDING_HIP_PARAMS_INIT \
    { (uint32_t)sizeof(unblending_hip_params), -1, 0.0, 0, 0, 0, 0, NULL, 0, 0, 0, 0, 0, 0, 0 }

typedef struct unblending_hip_stats {
    uint64_t evals;        /* objective+gradient evaluations, summed over pixels */
    uint64_t iterations;   /* search directions (two-loop recursions) */
    uint64_t outer;        /* augmented-Lagrangian iterations */
    uint64_t pixels;
    double solve_ms;       /* device time of the solve kernel(s) (timing = 1) */
    double filter_ms;      /* device time of the filter kernels (timing = 1) */
    double filter_coeff_ms;/* of which pass A */
    double filter_out_ms;  /* of which pass B */
    int32_t filter_radius; /* radius used */
    int32_t filter_channels;
    int32_t filter_launches;
    int32_t reserved;
    uint64_t fallback_pixels; /* warm start: pixels solved again from the default initialisation */
    double predict_ms;        /* warm start: device time of the model's prediction (timing = 1) */
    uint64_t lane_slots;      /* persistent solve: lane evaluations, useful or idle (evals / lane_slots
                                 is the share of lanes doing useful work) */
    uint64_t history_pairs;   /* (s, y) pairs the search directions' two-loop recursions went through */
} unblending_hip_stats;

UNBLENDING_API int32_t unblending_hip_unmix_ex(
    const unblending_rgb_view* image, const unblending_layer* layers, size_t layer_count,
    const unblending_options* options, const unblending_hip_params* params,
    double* output, size_t output_capacity, unblending_hip_stats* stats);

UNBLENDING_API int32_t unblending_hip_refine_ex(
    const unblending_rgb_view* image, const unblending_layer* layers, size_t layer_count,
    const double* first_pass, size_t first_pass_size, const unblending_options* options,
    const unblending_hip_params* params, double* output, size_t output_capacity,
    unblending_hip_stats* stats);

/* Bytes of device workspace the device-resident calls need with default
 * params (warm_start: kept for callers; the prediction needs none).  The
 * default inner minimiser keeps its (s, y) history in the workspace: up to
 * 1024 x 64 lanes x 1000 pairs x (8 L + 2) doubles, fewer blocks for images
 * under 65536 pixels. */
UNBLENDING_API size_t unblending_hip_workspace_size(int32_t width, int32_t height,
                                                   size_t layer_count, int32_t refine,
                                                   int32_t smooth_background, int32_t warm_start);
/* The same for the given params (NULL: defaults); 0 for invalid arguments. */
UNBLENDING_API size_t unblending_hip_workspace_size_ex(int32_t width, int32_t height,
                                                      size_t layer_count, int32_t refine,
                                                      int32_t smooth_background,
                                                      const unblending_hip_params* params);

UNBLENDING_API int32_t unblending_hip_unmix_device(
    const double* rgb, int32_t width, int32_t height, const unblending_layer* layers,
    size_t layer_count, const unblending_options* options, const unblending_hip_params* params,
    double* output, double* alpha_planes, void* workspace, size_t workspace_size, void* stream,
    unblending_hip_stats* stats);

UNBLENDING_API int32_t unblending_hip_refine_device(
    const double* rgb, int32_t width, int32_t height, const unblending_layer* layers,
    size_t layer_count, const double* first_pass, const double* alpha_planes,
    const unblending_options* options, const unblending_hip_params* params, double* output,
    void* workspace, size_t workspace_size, void* stream, unblending_hip_stats* stats);

/* Guided filter alone, device pointers: q[c] = filter(p[c]) for `channels`
 * planes p (each [height][width]) with RGB guidance rgb [height][width][3].
 * block / rows: launch shape as in unblending_hip_params (0: defaults).
 * Exposed for the filter's parity tests, bandwidth measurement and the
 * launch-shape autotuner. */
UNBLENDING_API int32_t unblending_hip_guided_filter_device(
    const double* rgb, int32_t width, int32_t height, const double* const* p, double* const* q,
    int32_t channels, int32_t radius, double epsilon, int32_t block, int32_t rows, void* workspace,
    size_t workspace_size, void* stream, unblending_hip_stats* stats);

UNBLENDING_API size_t unblending_hip_filter_workspace_size(int32_t width, int32_t height,
                                                          int32_t channels);

/* The guided filter without a workspace argument and without any device
 * allocation: no counterpart in the reference.  For callers that cannot
 * allocate -- code under HIP graph capture, serving loops that forbid
 * hipMalloc.  Results equal unblending_hip_guided_filter_device's for the same
 * arguments.  The filter's scratch (8 (4 channels width height + 32) bytes)
 * lives in a device array of 256 MiB that the library holds on each GPU it
 * runs on until the process ends; larger calls return
 * UNBLENDING_INVALID_ARGUMENT (unblending_hip_last_error says why).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU race.
 * stats (may be NULL) receives the call's shape, no device times. */
UNBLENDING_API int32_t unblending_hip_guided_filter_device_noalloc(
    const double* rgb, int32_t width, int32_t height, const double* const* p, double* const* q,
    int32_t channels, int32_t radius, double epsilon, int32_t block, int32_t rows, void* stream,
    unblending_hip_stats* stats);

/* ---- warm-start model (no counterpart in the reference) ----
 * A 27 -> 128 -> 128 -> 128 -> outputs network (GELU, sigmoid outputs) whose
 * input is a pixel's 3x3 RGB neighbourhood and whose outputs are the packed
 * first-pass values (alpha, R, G, B per layer; outputs = 4 * layers <= 32).
 * BF16 matrix-core products with FP32 accumulation; FP32 master weights and
 * Adam moments, all in one device buffer of unblending_hip_mlp_model_size()
 * bytes.  Parameter vector (FP32, padded): W1 [32][128], b1, W2 [128][128],
 * b2, W3 [128][128], b3, W4 [128][32], b4; weights [in][out]. */
UNBLENDING_API size_t unblending_hip_mlp_model_size(void);
UNBLENDING_API int32_t unblending_hip_mlp_param_count(void);  /* padded FP32 parameters */
/* Uploads FP32 parameters, zeroes the Adam moments, makes the BF16 copies. */
UNBLENDING_API int32_t unblending_hip_mlp_set_params(void* model, const float* params, void* stream);
UNBLENDING_API int32_t unblending_hip_mlp_get_params(const void* model, float* params, void* stream);
UNBLENDING_API size_t unblending_hip_mlp_train_workspace_size(int32_t batch);
/* Adam steps on samples drawn from `count` images [count][height][width][3]
 * with their solved first passes [count][layers][height*width][RGBA] as
 * targets.  steps_done is the number of steps already taken (bias
 * correction).  losses (host, may be NULL) receives one loss per step.
 * fixed_samples (device, may be NULL): batch pixel indices for every step. */
UNBLENDING_API int32_t unblending_hip_mlp_train(
    void* model, const double* images, int32_t count, int32_t width, int32_t height,
    const double* targets, int32_t layers, int32_t steps, int32_t steps_done, int32_t batch,
    float learning_rate, uint64_t seed, const int64_t* fixed_samples, int32_t update,
    void* workspace, size_t workspace_size, void* stream, float* losses, float* gradients);
/* Predicted first pass for one image, packed [layers][height*width][RGBA].
 * One fused kernel: no workspace is needed (the size is 0; workspace may be
 * NULL), the arguments are kept for callers that pass one. */
UNBLENDING_API size_t unblending_hip_mlp_predict_workspace_size(void);
UNBLENDING_API int32_t unblending_hip_mlp_predict(
    const void* model, const double* rgb, int32_t width, int32_t height, int32_t layers,
    double* output, void* workspace, size_t workspace_size, void* stream);

/* ---- batches of images (no counterpart in the reference, which solves one
 * image a call, unblending/sdk/unblending_c.cpp:50-110) ----
 * Device-resident images of any sizes, each with its own layer set (1 to 8
 * layers).  All images of one layer count are solved by one persistent
 * launch whose waves take pixel chunks of those images from a device queue;
 * the per-pixel results are those of unblending_hip_unmix_device /
 * unblending_hip_refine_device on each image alone.  The options apply to
 * every image.  Enqueued on `stream`; no allocation, no synchronisation
 * unless stats are asked for, so a call can be captured in a HIP graph.  The
 * refinement filters up to four images at once on streams of the library's
 * own, forked from and joined to `stream` by events (a graph capture records
 * them as branches); they are made at the first batch refinement on a device
 * outside a capture and kept (a refinement captured before then filters on
 * `stream` alone).
 * With params.warm_start = 1, the first pass starts each image that has a
 * model (unblending_hip_image.mlp) from that model's prediction, as
 * unblending_hip_unmix_device does for one image (params.mlp is not used). */
typedef struct unblending_hip_image {
    const double* rgb;               /* device [height][width][3] */
    int32_t width, height;
    const unblending_layer* layers;  /* host */
    size_t layer_count;
    const double* first_pass;        /* refinement: device packed first pass */
    double* alpha_planes;            /* first pass: optional output alpha planes [L][h][w];
                                        refinement: optional alpha planes of the first pass */
    double* output;                  /* device packed [layer][height][width][RGBA] */
    int32_t filter_radius;           /* refinement: < 0: 60 * min(width, height) / 1000 */
    int32_t reserved;
    const void* mlp;                 /* first pass with params.warm_start = 1: the trained model of this
                                        image's layer set (unblending_hip_mlp_*), or NULL for the
                                        default initialisation */
} unblending_hip_image;

UNBLENDING_API size_t unblending_hip_batch_workspace_size(const unblending_hip_image* images, size_t count,
                                                         int32_t refine, int32_t smooth_background);
UNBLENDING_API size_t unblending_hip_batch_workspace_size_ex(const unblending_hip_image* images, size_t count,
                                                            int32_t refine, int32_t smooth_background,
                                                            const unblending_hip_params* params);
UNBLENDING_API int32_t unblending_hip_unmix_batch_device(
    const unblending_hip_image* images, size_t count, const unblending_options* options,
    const unblending_hip_params* params, void* workspace, size_t workspace_size, void* stream,
    unblending_hip_stats* stats);
UNBLENDING_API int32_t unblending_hip_refine_batch_device(
    const unblending_hip_image* images, size_t count, const unblending_options* options,
    const unblending_hip_params* params, void* workspace, size_t workspace_size, void* stream,
    unblending_hip_stats* stats);

/* ---- image-level recomposite ----
 * Replaces the reference's CPU composite_layers(layers, comp_ops, modes)
 * (unblending/include/unblending/unblending.hpp:34-36,
 * unblending/src/unblending.cpp:638-664): the packed straight-alpha layers
 * [layer_count][height][width][RGBA] (bottom layer first) folded bottom to
 * top with each layer's Porter-Duff factors comp_ops[3 l .. 3 l + 2]
 * (X, Y, Z) and blend mode modes[l] (equations.cpp:10-74, no crop), into
 * output [height][width][RGBA].  1 to 16 layers.  The _device form takes
 * device pointers and is enqueued on `stream`; the host form copies in and
 * out and returns after the result is written. */
UNBLENDING_API int32_t unblending_hip_composite_device(
    const double* layers, int32_t width, int32_t height, size_t layer_count, const int32_t* comp_ops,
    const int32_t* modes, double* output, void* stream);
UNBLENDING_API int32_t unblending_hip_composite(
    const double* layers, size_t layers_size, int32_t width, int32_t height, size_t layer_count,
    const int32_t* comp_ops, const int32_t* modes, double* output, size_t output_capacity);

/* Planar seam, host pointers (see above).  Returns an unblending status. */
UNBLENDING_API int32_t uhip_solve_planes(
    int32_t width, int32_t height, const unblending_layer* layers, size_t layer_count,
    const double* rgb_planes, const double* first_planes, int32_t opaque, int32_t smooth,
    double* out_planes);

/* Last error message of the calling thread (empty when none). */
UNBLENDING_API const char* unblending_hip_last_error(void);

/* No counterpart in the reference.  The host-pointer entries
 * (unblending_unmix, unblending_refine and their _ex forms, uhip_solve_planes,
 * unblending_hip_composite) keep their device buffers -- image, first pass,
 * result and workspace, the largest up to ~31 GB for a 4096^2 7-layer image --
 * on each device between calls, growing them when a call needs more, because
 * allocating them again at every call makes the driver clear the memory each
 * time.  This frees the buffers kept on the current device (after a call in
 * progress on another thread).  Returns UNBLENDING_OK. */
UNBLENDING_API int32_t unblending_hip_release_buffers(void);

#ifdef __cplusplus
}
#endif
#endif
