// sht_noalloc.hip -- the fixed device workspace behind cu_SH_to_spat_batch_noalloc.
//
// `scratch` is a file-scope device array: each GPU the library runs on holds its own copy,
// at its own address, from module load until the process ends.  In host code the name is
// only a handle for the runtime, so the device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and only that address is handed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sht_internal.h"

static __device__ __align__(256) char scratch[shtns_amd::kNoallocScratchBytes];

namespace shtns_amd {

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;
}  // namespace

int noalloc_scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (g_scratch_ptr[dev]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        return -1;
    }
    g_scratch_ptr[dev] = p;
    return 0;
}

void *noalloc_scratch_ptr(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    return g_scratch_ptr[dev];
}

}  // namespace shtns_amd
