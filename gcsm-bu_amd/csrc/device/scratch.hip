// Library-owned static partial-result ring for csm_apply_range_static.
//
// `scratch` is a file-scope device array: each GPU the library runs on holds its own copy for the
// life of the process, at its own address.  Host code never passes the array's name to a kernel
// (on the host it is only a stand-in symbol); static_ring() resolves the device address once per
// device with hipGetSymbolAddress, with that device current, and every later use goes through the
// returned pointer.  Calls that use the array on one GPU share it, so overlapping calls race.
#include "kernels.h"

#include <mutex>

namespace csm {

static __device__ __align__(256) uint32_t scratch[kStaticRingWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t static_ring(int device, uint32_t** ptr) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current first
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = static_cast<uint32_t*>(p);
    }
    *ptr = g_ptr[device];
    return hipSuccess;
}

}  // namespace csm
