// Fixed device scratch for the allocation-free H_AC entry point (imps_hac_apply_static): one
// file-scope device array per GPU, resolved to its device address once per device.  The array's
// name is a host-side stand-in; only the address from hipGetSymbolAddress is used by host code.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "static_scratch.hpp"

namespace imps {

static __device__ __align__(256) char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

void* static_scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw InvalidArg("static scratch: device index out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[device]) {
        int cur = 0;
        hip_check(hipGetDevice(&cur), "hipGetDevice");
        hip_check(hipSetDevice(device), "hipSetDevice");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

}  // namespace imps
