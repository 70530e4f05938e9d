// ub_noalloc.hip -- the device scratch of the allocation-free guided-filter
// entry point (unblending_hip_guided_filter_device_noalloc).
//
// The filter's scratch -- the bounds slots and every channel's coefficient
// planes, 8 (4 C n + 32) bytes for C channels of n pixels -- lives in a
// file-scope device array of kNoAllocScratchBytes, which each GPU the library
// runs on holds until the process ends.  Host code uses the array only through
// the device address hipGetSymbolAddress returns for it, looked up once per
// device with that device current; the array's name is a host-side stand-in
// and never reaches a kernel.  Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU race.
#include <mutex>

#include "ub_host.hpp"

namespace ubhip {
namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(void** ptr)
{
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace ubhip
