// ub_static.hip -- the fixed frame scratch of ubg_render_noalloc.
//
// One file-scope device array, `scratch`, of UBG_NOALLOC_SCRATCH_BYTES: every
// GPU the library runs on holds its own copy (allocated with the code object,
// kept until the process ends). Host code never uses the array's name as an
// address -- on the host it is only a stand-in symbol -- but the device
// address that hipGetSymbolAddress returns for it, looked up once per device
// with that device current and kept here per device.

#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/umbreon_gpu.h"
#include "ub_kernels.h"

__device__ __align__(256) unsigned char scratch[UBG_NOALLOC_SCRATCH_BYTES];

namespace ubg {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_looked[kMaxDevices] = {};
}  // namespace

hipError_t staticScratchAddress(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_looked[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
    g_status[device] = e;
    g_addr[device] = e == hipSuccess ? p : nullptr;
    g_looked[device] = true;
  }
  *out = g_addr[device];
  return g_status[device];
}

}  // namespace ubg
