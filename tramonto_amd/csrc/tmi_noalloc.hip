// Device scratch of the allocation-free entry point (tmi_fill_residual_noalloc): a file-scope device
// array, one per GPU the library runs on, held until the process ends.  Host code never uses the
// array's name as an address: at set-up (problem creation, or the entry point's first call on a
// device) its device address is looked up once per device with hipGetSymbolAddress, with that device
// current, and kept per device; kernels receive that pointer.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>

#include "tmi_internal.hpp"

namespace tmi {
namespace {

constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kNoAllocScratchDoubles];

double *g_ptr[kMaxDevices] = {};
std::mutex g_mu;

}  // namespace

double *noalloc_scratch()
{
    int dev = 0;
    TMI_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::domain_error("device index beyond the library's scratch table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        TMI_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_ptr[dev] = (double *)p;
    }
    return g_ptr[dev];
}

}  // namespace tmi
