// noalloc.hip -- the fixed device scratch array of the allocation-free entry point
// (ndg_record_receivers_noalloc) and its per-device address.
//
// The array is a file-scope __device__ variable: every GPU the library runs on holds its own copy
// for the life of the process. Host code never uses the array's name as an address (on the host
// it only stands for the symbol); it asks the runtime for the device address once per device, with
// that device current, and passes that address to kernels and copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

static __device__ __align__(256) double scratch[ndg::kNoAllocScratchBytes / sizeof(double)];

namespace ndg {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(double **ptr) {
    *ptr = nullptr;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = (double *)p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace ndg
