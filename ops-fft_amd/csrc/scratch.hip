// scratch.hip - the device-resident scratch area of the allocation-free execution entry point
// (aclfftExtExecC2CNoAlloc).
//
// `scratch` is a file-scope device array: every GPU the library runs on holds its own copy, at its
// own address, from the module's load until the process ends.  In host code the array's name is
// only a stand-in; the device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and kept, and only that address is handed to kernels and copies.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "ofx_launch.h"

namespace {

static __device__ __align__(256) unsigned char scratch[OFX_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::atomic<void*> g_ptr[kMaxDevices];
std::mutex g_mu;

}  // namespace

extern "C" int ofx_scratch(void** ptr, size_t* bytes) {
    if (ptr == nullptr) return (int)hipErrorInvalidValue;
    *ptr = nullptr;
    if (bytes) *bytes = OFX_SCRATCH_BYTES;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
    void* p = g_ptr[dev].load(std::memory_order_acquire);
    if (p == nullptr) {
        std::lock_guard<std::mutex> lk(g_mu);
        p = g_ptr[dev].load(std::memory_order_relaxed);
        if (p == nullptr) {
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || p == nullptr) {
                (void)hipGetLastError();
                return (int)(e != hipSuccess ? e : hipErrorInvalidSymbol);
            }
            g_ptr[dev].store(p, std::memory_order_release);
        }
    }
    *ptr = p;
    return 0;
}
