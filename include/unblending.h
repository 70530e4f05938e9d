/*
 * unblending.h -- versioned C interface of the two Unblending passes.
 *
 * ABI-compatible with the reference's unblending/include/unblending.h (v1):
 * the same error codes, blend-mode ids, structs and the three entry points
 * below, so an existing C client relinks against this library unchanged.
 * The one addition is UNBLENDING_BACKEND_HIP, the gfx950 backend this
 * library implements.
 *
 *   unblending_abi_version ... replaces unblending/include/unblending.h:85
 *                              (unblending/sdk/unblending_c.cpp:95)
 *   unblending_unmix ......... replaces unblending/include/unblending.h:92-94
 *                              (unblending/sdk/unblending_c.cpp:97-102)
 *   unblending_refine ........ replaces unblending/include/unblending.h:98-101
 *                              (unblending/sdk/unblending_c.cpp:104-110)
 *
 * Backends in this library: HIP and AUTO run on the current HIP device
 * (AUTO does not fall back to a CPU solver: this library ships none); a
 * missing device or a device failure returns UNBLENDING_DEVICE_ERROR.  CPU
 * returns UNBLENDING_INVALID_ARGUMENT and CUDA returns UNBLENDING_DEVICE_ERROR
 * (not built), as the reference does for a backend it was built without.
 */
#ifndef UNBLENDING_C_API_H
#define UNBLENDING_C_API_H

#include <stddef.h>
#include <stdint.h>

#define UNBLENDING_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

#define UNBLENDING_ABI_VERSION 1u

/* Status codes (unblending.h:23-30). */
enum {
    UNBLENDING_OK = 0,
    UNBLENDING_INVALID_ARGUMENT = 1,
    UNBLENDING_BUFFER_TOO_SMALL = 2,
    UNBLENDING_CUDA_ERROR = 3,        /* the device error code of ABI v1 */
    UNBLENDING_DEVICE_ERROR = 3,      /* the same code, named for any device */
    UNBLENDING_OUT_OF_MEMORY = 4,
    UNBLENDING_SOLVER_ERROR = 5
};

/* Backends (unblending.h:32-36), plus HIP. */
enum {
    UNBLENDING_BACKEND_AUTO = 0,
    UNBLENDING_BACKEND_CPU = 1,
    UNBLENDING_BACKEND_CUDA = 2,
    UNBLENDING_BACKEND_HIP = 3
};

/* Separable blend modes (unblending.h:38-52). */
enum {
    UNBLENDING_NORMAL = 0,
    UNBLENDING_MULTIPLY = 1,
    UNBLENDING_SCREEN = 2,
    UNBLENDING_OVERLAY = 3,
    UNBLENDING_DARKEN = 4,
    UNBLENDING_LIGHTEN = 5,
    UNBLENDING_COLOR_DODGE = 6,
    UNBLENDING_COLOR_BURN = 7,
    UNBLENDING_HARD_LIGHT = 8,
    UNBLENDING_SOFT_LIGHT = 9,
    UNBLENDING_DIFFERENCE = 10,
    UNBLENDING_EXCLUSION = 11,
    UNBLENDING_LINEAR_DODGE = 12
};

/* Row-major RGB triples in [0, 1]; size and row_stride count doubles; a zero
 * stride means rows of exactly 3 * width doubles. */
typedef struct unblending_rgb_view {
    const double* data;
    size_t size;
    int32_t width;
    int32_t height;
    size_t row_stride;
} unblending_rgb_view;

/* One layer, bottom to top.  Porter-Duff factors (1,1,1) = source-over,
 * (2,1,1) = plus; Gaussian colour model with a symmetric positive-definite
 * precision (inverse covariance) in row-major order. */
typedef struct unblending_layer {
    int32_t comp_x, comp_y, comp_z;
    int32_t blend_mode;
    double mean[3];
    double precision[9];
} unblending_layer;

typedef struct unblending_options {
    uint32_t struct_size;
    int32_t backend;
    int32_t opaque_background;
    int32_t smooth_background;
    int32_t target_concurrency;
} unblending_options;

#define UNBLENDING_OPTIONS_INIT \
    { (uint32_t)sizeof(unblending_options), UNBLENDING_BACKEND_AUTO, 1, 0, 0 }

UNBLENDING_API uint32_t unblending_abi_version(void);

/* First pass.  Output: caller-owned packed [layer][row][column][RGBA]
 * straight-alpha doubles, capacity >= 4 * width * height * layer_count;
 * written only on success.  NULL options select UNBLENDING_OPTIONS_INIT. */
UNBLENDING_API int32_t unblending_unmix(
    const unblending_rgb_view* image, const unblending_layer* layers, size_t layer_count,
    const unblending_options* options, double* output, size_t output_capacity);

/* Second pass.  first_pass holds exactly 4 * width * height * layer_count
 * doubles and may alias output. */
UNBLENDING_API int32_t unblending_refine(
    const unblending_rgb_view* image, const unblending_layer* layers, size_t layer_count,
    const double* first_pass, size_t first_pass_size, const unblending_options* options,
    double* output, size_t output_capacity);

#ifdef __cplusplus
}
#endif
#endif
