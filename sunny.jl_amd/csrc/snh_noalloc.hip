// Static device scratch for the entry point that allocates nothing
// (snh_energy_noalloc): the per-workgroup FP64 partial energy sums live in a
// file-scope device array instead of a buffer the library allocates.  Also
// the allocation-free bookkeeping of dynamic-LDS limits its launches use.
//
// Each GPU holds its own copy of the array, at its own address, until the
// process ends.  Host code never passes the array's name to a kernel (in host
// code it is a stand-in, not a device address): snh_create resolves the
// address of the current device's copy once with hipGetSymbolAddress and
// keeps it per device; kernels receive that pointer.
#include <mutex>

#include "snh_internal.h"

namespace snh {

static __device__ __align__(256) double scratch[kNoAllocPartials];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch_init()
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return hipSuccess;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
    return hipSuccess;
}

double* noalloc_scratch(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[dev];
}

void allow_dynamic_lds(const void* fn, size_t bytes)
{
    static std::mutex mu;
    static const void* fns[256];
    static size_t have[256];
    static int n = 0;
    std::lock_guard<std::mutex> lock(mu);
    int k = 0;
    while (k < n && fns[k] != fn) k++;
    if (k < n && have[k] >= bytes) return;
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (k == n && n < 256) fns[n++] = fn;
    if (k < n) have[k] = bytes;
}

}  // namespace snh
