// ub_composite.hip -- image-level recomposite of a packed layer stack on gfx950.
//
// The device counterpart of the reference's CPU composite_layers
// (unblending/src/unblending.cpp:638-664): per pixel, the layers'
// straight-alpha RGBA folded bottom to top with the Porter-Duff composite of
// unblending/src/equations.cpp:10-74 (no crop), every expression in the
// order the solve kernels and the host oracle use.  One pixel per lane; the
// layer table (comp-op factors and blend mode) is wave-uniform.
#include "ub_kernels.hpp"

namespace ubhip {
namespace {

__global__ void __launch_bounds__(256) composite_kernel(CompositeArgs a)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double* p0 = a.layers + 4 * i;
    double acc[3] = {p0[0], p0[1], p0[2]};
    double accA = p0[3];
    for (int k = 1; k < a.count; ++k) {
        const CompositeLayer l = a.table[k];
        const int mode = __builtin_amdgcn_readfirstlane(l.mode);
        const double* pk = a.layers + 4 * ((long long)k * a.n + i);
        const double cs[3] = {pk[0], pk[1], pk[2]};
        const double as = pk[3], ad = accA;
        const double A = l.X * as * ad + l.Y * as * (1.0 - ad) + l.Z * ad * (1.0 - as);
        const double ky = l.Y * as * (1.0 - ad);
        const double kz = l.Z * ad * (1.0 - as);
        double f[3];
        blend3(mode, cs, acc, f);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double p = f[c] * as * ad + ky * cs[c] + kz * acc[c];
            acc[c] = (A > 1e-12) ? p / A : p;
        }
        accA = A;
    }
    double* o = a.out + 4 * i;
    o[0] = acc[0];
    o[1] = acc[1];
    o[2] = acc[2];
    o[3] = accA;
}

}  // namespace

hipError_t launch_composite(const CompositeArgs& args, hipStream_t stream)
{
    if (args.n <= 0) return hipSuccess;
    if (args.count < 1 || args.count > kMaxCompositeLayers) return hipErrorInvalidValue;
    hipLaunchKernelGGL(composite_kernel, dim3((unsigned)((args.n + 255) / 256)), dim3(256), 0, stream, args);
    return hipGetLastError();
}

}  // namespace ubhip
