// ub_solve_var.hip -- launches of the persistent solve kernels
// (ub_solve_impl.hpp) for the reference's problem variants: layer 1's colour
// constrained to gray (SECOND_LAYER_GRAY, unblending.cpp:21,168-172) and the
// first pass's sparsity term (SPARCITY, unblending.cpp:17,236-240).  They are
// compiled for the history-keeping kernel (the default minimiser) only, in a
// source file of their own so that they build beside the plain kernels.
#include "ub_solve_impl.hpp"

namespace ubhip {
namespace {

template <int L, int VAR>
hipError_t launch_var(bool refine, const BatchArgs& args, hipStream_t stream)
{
    if constexpr ((VAR & kVarGray) != 0 && L < 2) {
        return hipErrorInvalidValue;
    } else {
        if (refine) {
            // the sparsity term is a first-pass term (unblending.cpp:236-240)
            if constexpr ((VAR & kVarGray) == 0) return hipErrorInvalidValue;
            else return launch_batch<L, true, true, kVarGray>(args, stream);
        }
        return launch_batch<L, false, true, VAR>(args, stream);
    }
}

template <int L>
hipError_t launch_var_L(bool refine, int var, const BatchArgs& args, hipStream_t stream)
{
    switch (var) {
    case kVarGray: return launch_var<L, kVarGray>(refine, args, stream);
    case kVarSparse: return launch_var<L, kVarSparse>(refine, args, stream);
    case kVarGray | kVarSparse: return launch_var<L, kVarGray | kVarSparse>(refine, args, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_batch_variant(int L, bool refine, const BatchArgs& args, hipStream_t stream)
{
    if (args.sp.memory == 1 || !args.hist || args.hist_cap < 1 || args.hist_blocks < 1) return hipErrorInvalidValue;
    const int var = (args.sp.gray ? kVarGray : 0) | (!refine && args.sp.sparse_sigma != 0.0 ? kVarSparse : 0);
    switch (L) {
    case 1: return launch_var_L<1>(refine, var, args, stream);
    case 2: return launch_var_L<2>(refine, var, args, stream);
    case 3: return launch_var_L<3>(refine, var, args, stream);
    case 4: return launch_var_L<4>(refine, var, args, stream);
    case 5: return launch_var_L<5>(refine, var, args, stream);
    case 6: return launch_var_L<6>(refine, var, args, stream);
    case 7: return launch_var_L<7>(refine, var, args, stream);
    case 8: return launch_var_L<8>(refine, var, args, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace ubhip
