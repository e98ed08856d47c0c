// dq_noalloc.hip -- the fixed device workspace of dq_spectrum_noalloc.
//
// One file-scope device array per GPU: the spectrum's windowed C(n) buffer
// and the FFT's ping-pong buffer, 2 x nfft complex doubles, nfft at most
// kNoAllocMaxNfft.  Host code never passes `scratch` itself to a kernel or a
// copy (in host code the name is a host-side stand-in, not a device address):
// the address of each device's copy comes from hipGetSymbolAddress, looked up
// once per device with that device current, and kept in `g_addr`.
#include <mutex>

#include "dq_internal.h"

namespace dq {

static __device__ __align__(256) double2 scratch[2 * kNoAllocMaxNfft];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double2* g_addr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(double2** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_addr[dev] = static_cast<double2*>(p);
    }
    *out = g_addr[dev];
    return hipSuccess;
}

}  // namespace dq
