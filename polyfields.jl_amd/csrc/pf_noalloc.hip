// Device scratch for the allocation-free entry point (pf_scfstress_noalloc): a file-scope
// array of fixed size in every GPU's copy of the code object, so the call needs no
// hipMalloc.  Host code never passes the array's name to a kernel or a copy (on the host
// it names a stand-in, not a device address): each device's address is looked up once,
// with that device current, and kept here.
#include <mutex>

#include "pf_host.h"

namespace pf {

namespace {
constexpr size_t kScratchBytes = size_t(32) << 20;   // 32 MiB per GPU
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
char* scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

// The device address of `scratch` on `device` (made current by the caller) and its size.
char* noalloc_scratch(int device, size_t* bytes) {
    if (device < 0 || device >= kMaxDevices) throw Error(-1, "device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        int cur = -1;
        PF_HIP(hipGetDevice(&cur));
        if (cur != device) throw Error(-2, "noalloc_scratch: the device is not current");
        void* p = nullptr;
        PF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(-3, "hipGetSymbolAddress returned no address for the scratch array");
        scratch_ptr[device] = static_cast<char*>(p);
    }
    *bytes = kScratchBytes;
    return scratch_ptr[device];
}

}  // namespace pf
