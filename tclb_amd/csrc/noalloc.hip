// noalloc.hip -- the device staging memory of tclbx_get_quantity_noalloc, the
// one entry point of the library that allocates nothing: its quantity kernel
// writes into a file-scope device array, which the library reserves on each
// GPU it runs on for the life of the process, and the result is copied from
// there to the caller's host buffer.
#include <mutex>
#include <stdexcept>
#include <vector>

#include "lattice.h"

namespace tclbx {
namespace {

// 64 MiB of doubles: a scalar quantity of up to 8,388,608 nodes, a vector
// quantity of up to 2,796,202
constexpr size_t kScratchDoubles = size_t(8) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
std::vector<double*> g_ptr;  // device address of `scratch` on each device

}  // namespace

size_t quantity_scratch_doubles() { return kScratchDoubles; }

// The address of the current device's copy of `scratch`, looked up once per
// device with that device current.  The array's name in host code is a
// host-side stand-in, never a device address: kernels get this pointer.
double* quantity_scratch() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_ptr.size() <= dev) g_ptr.resize(dev + 1, nullptr);
    if (g_ptr[dev] == nullptr) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (p == nullptr) throw std::runtime_error("hipGetSymbolAddress(scratch) gave a null address");
        g_ptr[dev] = (double*)p;
    }
    return g_ptr[dev];
}

}  // namespace tclbx
