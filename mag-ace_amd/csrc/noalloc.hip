// noalloc.hip -- fixed device scratch for magace_compute_device_noalloc.
//
// The evaluation keeps per-atom intermediates between its three kernels: the feature vector
// Z and its adjoint W (complex, nz entries per atom each) and five per-atom scalars.
// magace_compute_device takes them from buffers the library grows with hipMalloc;
// magace_compute_device_noalloc takes them from the file-scope array below instead, so that
// callers that must not allocate (code under graph capture, tight serving loops) can
// evaluate.  The array lives on every device the library's code object is loaded on, until
// the process ends.  Host code must not pass `scratch` itself to a kernel: on the host the
// name is a stand-in, not a device address.  The address is looked up once per device with
// hipGetSymbolAddress (scratch_setup, run when the potential's tables are uploaded, with that
// device current) and only that pointer is used afterwards.
#include <hip/hip_runtime.h>

#include <mutex>

#include "magace.hpp"

namespace magace {
namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;  // 128 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

void* g_scratch[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

}  // namespace

int scratch_setup() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return -1;
    g_scratch[dev] = p;
    return 0;
}

void* scratch_device_ptr(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[dev];
}

size_t scratch_bytes() { return kScratchBytes; }

}  // namespace magace
