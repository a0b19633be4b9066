// pbf_static.hip - the library's static device scratch: every temporary of
// pbf_step_noalloc lives in one file-scope device array, so that call
// allocates no device memory.  Each GPU that loads the library holds its own
// copy of the array, from module load to process end.
#include <mutex>

#include "pbf_kernels.hpp"

static __device__ __align__(256) unsigned char
    scratch[pbf::kStaticScratchBytes];

namespace pbf {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of the calling thread's current device's copy of the
// scratch array (looked up once per device, with that device current; the
// array's host-side name is no device address).
hipError_t static_scratch(void **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace pbf
