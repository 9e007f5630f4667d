// gbsw_noalloc.hip -- device scratch of the allocation-free entry point
// (gbsw_align_noalloc, include/gbsw.h).
//
// The scratch is one file-scope device array, so each device the library runs
// on holds its own copy (kNoallocScratchBytes) until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory: the device
// address of a device's copy comes from hipGetSymbolAddress, taken once per
// device at set-up (gbsw_context_create, with that device current) and kept
// here; every use of the scratch goes through that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gbsw_internal.hpp"

namespace gbsw {

static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};
} // namespace

int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(GBSW_E_ARGUMENT, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return GBSW_OK;
    int cur = -1;
    GBSW_HIP_TRY(hipGetDevice(&cur));
    if (cur != device) return fail(GBSW_E_INTERNAL, "scratch set-up needs its device current");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(GBSW_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return GBSW_OK;
}

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_scratch[device];
}

} // namespace gbsw
