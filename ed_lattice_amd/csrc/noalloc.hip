// Device memory for the allocation-free operator entry point
// (edl_hxv_noalloc): the two-pass intermediate T lives in one file-scope
// device array of kNoAllocScratchBytes, which every GPU the library runs on
// holds until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is a stand-in, not a device address): each
// device's address is looked up once with hipGetSymbolAddress when the
// first context on that device is created, and only that pointer is used.
#include <mutex>

#include "edl_internal.hpp"

namespace edl {

constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;  // 256 MiB
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_noalloc_mutex;
static void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// With `device` current: record that device's address of `scratch` (once).
void noalloc_scratch_setup(int device) {
  EDL_CHECK(device >= 0 && device < kNoAllocMaxDevices, "noalloc_scratch_setup: device index beyond 63");
  std::lock_guard<std::mutex> g(g_noalloc_mutex);
  if (g_noalloc_ptr[device]) return;
  void* p = nullptr;
  EDL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  EDL_CHECK(p != nullptr, "noalloc_scratch_setup: no device address for the scratch array");
  g_noalloc_ptr[device] = p;
}

// The device address of `scratch` on `device` (set up by edl_ctx_create).
void* noalloc_scratch(int device) {
  EDL_CHECK(device >= 0 && device < kNoAllocMaxDevices, "noalloc_scratch: device index beyond 63");
  std::lock_guard<std::mutex> g(g_noalloc_mutex);
  EDL_CHECK(g_noalloc_ptr[device] != nullptr, "noalloc_scratch: scratch not set up on this device");
  return g_noalloc_ptr[device];
}

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

}  // namespace edl
