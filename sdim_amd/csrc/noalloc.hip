// Device memory of the allocation-free entry point (sdim_simulate_noalloc): the
// metadata of its submissions (item descriptors, launch order, lowered ops, noise
// thresholds, launch queues, counters) lives in `scratch`, a file-scope device array
// of SDIM_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
// process ends.  Host code never uses the array's name as an address: at context
// creation (with the device current) hipGetSymbolAddress yields the array's address
// on that device, which is kept per device and handed to the kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "engine.h"

namespace {

static __device__ __align__(256) uint8_t scratch[SDIM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};   // address of `scratch` on each device

}  // namespace

namespace sdim {

int noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return 0;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
  g_ptr[device] = (uint8_t*)p;
  return 0;
}

uint8_t* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace sdim
