// olap_noalloc.hip -- device scratch of olap_rowwise_noalloc.
//
// The allocation-free row-wise entry point keeps all of its device scratch
// (dimension statistics, the error word, the total count and the packed
// dimension vectors) in one file-scope array, so a call allocates nothing.
// Each GPU holds its own copy of the array for the life of the process. The
// array's name is a host-side stand-in: host code uses only the device
// address that hipGetSymbolAddress returns for the current device, taken
// once per device when a context is created there.
#include <hip/hip_runtime.h>

#include <mutex>

#include "olap_internal.h"

static __device__ __align__(256) unsigned char scratch[olap::kNoallocScratchBytes];

namespace olap {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices];
}  // namespace

hipError_t noalloc_scratch_init(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the caller makes it current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_addr[device] = p;
    }
    *ptr = g_addr[device];
    return hipSuccess;
}

}  // namespace olap
