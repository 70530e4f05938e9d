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
    int32_t reserved;
} unblending_hip_params;

#define UNBLENDING_HIP_PARAMS_INIT \
    { (uint32_t)sizeof(unblending_hip_params), -1, 0.0, 0, 0, 0, 0 }

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

/* Bytes of device workspace the device-resident calls need. */
UNBLENDING_API size_t unblending_hip_workspace_size(int32_t width, int32_t height,
                                                   size_t layer_count, int32_t refine,
                                                   int32_t smooth_background);

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
 * Exposed for the filter's parity tests and bandwidth measurement. */
UNBLENDING_API int32_t unblending_hip_guided_filter_device(
    const double* rgb, int32_t width, int32_t height, const double* const* p, double* const* q,
    int32_t channels, int32_t radius, double epsilon, void* workspace, size_t workspace_size,
    void* stream, unblending_hip_stats* stats);

UNBLENDING_API size_t unblending_hip_filter_workspace_size(int32_t width, int32_t height,
                                                          int32_t channels);

/* Planar seam, host pointers (see above).  Returns an unblending status. */
UNBLENDING_API int32_t uhip_solve_planes(
    int32_t width, int32_t height, const unblending_layer* layers, size_t layer_count,
    const double* rgb_planes, const double* first_planes, int32_t opaque, int32_t smooth,
    double* out_planes);

/* Last error message of the calling thread (empty when none). */
UNBLENDING_API const char* unblending_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
