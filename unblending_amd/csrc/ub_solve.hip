// ub_solve.hip -- launches of the per-pixel solve kernels (ub_solve_impl.hpp)
// for the plain problem, and the small kernels of the solve's set-up.
#include "ub_solve_impl.hpp"

namespace ubhip {

// Writes one image descriptor (passed in the argument segment) to device
// memory, in stream order.
__global__ void store_desc_kernel(ImageDesc* dst, ImageDesc d)
{
    const int i = threadIdx.x;
    constexpr int kWords = sizeof(ImageDesc) / sizeof(unsigned long long);
    static_assert(sizeof(ImageDesc) % sizeof(unsigned long long) == 0, "descriptor size");
    const unsigned long long* src = reinterpret_cast<const unsigned long long*>(&d);
    unsigned long long* out = reinterpret_cast<unsigned long long*>(dst);
    for (int k = i; k < kWords; k += blockDim.x) out[k] = src[k];
}

hipError_t launch_store_desc(ImageDesc* dst, const ImageDesc& d, hipStream_t stream)
{
    hipLaunchKernelGGL(store_desc_kernel, dim3(1), dim3(64), 0, stream, dst, d);
    return hipGetLastError();
}

__global__ void zero_kernel(unsigned long long* p, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0ull;
}

hipError_t launch_zero(unsigned long long* p, int count, hipStream_t stream)
{
    hipLaunchKernelGGL(zero_kernel, dim3((count + 63) / 64), dim3(64), 0, stream, p, count);
    return hipGetLastError();
}

namespace {

template <int L>
hipError_t launch_batch_L(bool refine, const BatchArgs& args, hipStream_t stream)
{
    if (args.sp.memory != 1) {
        if (!args.hist || args.hist_cap < 1 || args.hist_blocks < 1) return hipErrorInvalidValue;
        return refine ? launch_batch<L, true, true>(args, stream) : launch_batch<L, false, true>(args, stream);
    }
    return refine ? launch_batch<L, true, false>(args, stream) : launch_batch<L, false, false>(args, stream);
}

}  // namespace

hipError_t launch_batch_solve(int L, bool refine, const BatchArgs& args, hipStream_t stream)
{
    if (args.chunk_count <= 0) return hipSuccess;
    // the problem variants: kernels of their own (ub_solve_var.hip)
    if (args.sp.gray || (!refine && args.sp.sparse_sigma != 0.0)) return launch_batch_variant(L, refine, args, stream);
    switch (L) {
    case 1: return launch_batch_L<1>(refine, args, stream);
    case 2: return launch_batch_L<2>(refine, args, stream);
    case 3: return launch_batch_L<3>(refine, args, stream);
    case 4: return launch_batch_L<4>(refine, args, stream);
    case 5: return launch_batch_L<5>(refine, args, stream);
    case 6: return launch_batch_L<6>(refine, args, stream);
    case 7: return launch_batch_L<7>(refine, args, stream);
    case 8: return launch_batch_L<8>(refine, args, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_solve(int L, bool refine, const PixelArgs& args, hipStream_t stream)
{
    if (args.n <= 0) return hipSuccess;
    switch (L) {
    case 1: return launch_L<1>(refine, args, stream);
    case 2: return launch_L<2>(refine, args, stream);
    case 3: return launch_L<3>(refine, args, stream);
    case 4: return launch_L<4>(refine, args, stream);
    case 5: return launch_L<5>(refine, args, stream);
    case 6: return launch_L<6>(refine, args, stream);
    case 7: return launch_L<7>(refine, args, stream);
    case 8: return launch_L<8>(refine, args, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ubhip
