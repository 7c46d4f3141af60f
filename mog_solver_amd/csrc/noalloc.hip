// Device-resident scratch of mog_step_noalloc (include/mog.h).
//
// The step's only scratch beyond the context's own buffers is the Morton
// visiting order of its gather (keys, indices and the radix sort's
// temporaries), which mog_step allocates with hipMalloc on its first call
// for a particle count.  Callers that cannot allocate (code under HIP graph
// capture, tight serving loops) get it from this file-scope array instead.
// In host code `scratch` names a host-side stand-in, not a device address:
// its device address on each GPU is looked up once per device with
// hipGetSymbolAddress and only that address is ever passed to kernels.
#include <mutex>

#include "mog_common.h"

namespace mog {

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(void** ptr) {
  *ptr = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  return hipSuccess;
}

}  // namespace mog
