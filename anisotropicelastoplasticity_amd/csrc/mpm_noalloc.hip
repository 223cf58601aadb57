// mpm_noalloc.hip -- the fixed device scratch of the allocation-free entry point
// (mpm_surrogate_gradient_noalloc): one file-scope array per GPU, its address taken once per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mpm_mlp.h"
#include "mpm_noalloc.h"

namespace mpm {

// Every GPU the library runs on holds this array (about 9.3 MB) until the process ends.
static __device__ __align__(256) float scratch[NOALLOC_SCRATCH_FLOATS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
float* g_ptr[MAX_DEVICES] = {};
}  // namespace

// The device address of `scratch` on `device`, which must be the current device. The name of the
// array is a host-side stand-in: only the address returned here may reach kernels and copies.
hipError_t noalloc_scratch(int device, float** ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = static_cast<float*>(p);
  }
  *ptr = g_ptr[device];
  return hipSuccess;
}

}  // namespace mpm
