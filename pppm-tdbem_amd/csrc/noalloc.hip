// Scratch memory of the allocation-free entry point (pppm_run_noalloc, include/pppm_hip.h).
//
// One file-scope device array of kScratchBytes; every GPU the library runs on holds its own
// copy until the process ends. Host code never uses the array's name as an address (on the
// host it is only a stand-in symbol): the device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and kept per device.
#include "internal.hpp"

#include <mutex>

static __device__ __align__(256) unsigned char scratch[pppm_amd::kScratchBytes];

namespace pppm_amd
{

namespace
{
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_ptr(void **out)
{
    *out = nullptr;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return e;
    if (dev < 0 || dev >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev])
    {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace pppm_amd
