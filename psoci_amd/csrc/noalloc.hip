// Device scratch of the allocation-free sigma (psoci_sigma_noalloc).
//
// The zero-padded column staging of a sigma call (pad.hpp) normally lives in
// context memory grown with hipMalloc.  Callers that cannot allocate (code
// under stream capture, tight serving loops, harnesses that forbid
// allocation) get it from this fixed array instead: every GPU the library
// runs on holds its own copy, kScratchBytes large, until the process ends.
// In host code the array's name is only a host-side stand-in, so its device
// address is taken once per device with hipGetSymbolAddress (device
// current) by psoci_ctx_create, and only that address is handed to kernels.
// Calls that overlap on one GPU share the array and race; nothing in it has
// to start at zero (every staged element is written before it is read).
#include <mutex>

#include "ctx.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_mu;
bool g_done[kMaxDevices] = {};
void* g_ptr[kMaxDevices] = {};
hipError_t g_rc[kMaxDevices] = {};

}  // namespace

namespace psoci {

// Once per device, with that device current: the address of this device's
// copy of the scratch array.
int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return PSOCI_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_done[device]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_rc[device] = rc;
    g_ptr[device] = rc == hipSuccess ? p : nullptr;
    if (rc != hipSuccess) (void)hipGetLastError();
    g_done[device] = true;
  }
  return g_rc[device] == hipSuccess ? PSOCI_OK : PSOCI_ERR_HIP;
}

void noalloc_scratch(int device, void** p, size_t* bytes) {
  PSOCI_CHECK(device >= 0 && device < kMaxDevices, PSOCI_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  PSOCI_CHECK(g_done[device] && g_ptr[device], PSOCI_ERR_HIP, "the scratch array of this device has no address");
  *p = g_ptr[device];
  *bytes = kScratchBytes;
}

}  // namespace psoci
