// Staging memory of the allocation-free download (gdem_scene_download_noalloc).
//
// The download gathers a scene's state from its internal (sorted) order into
// original particle order before the copy to the host.  gdem_scene_download
// allocates that staging per call; callers that must not allocate device
// memory (code between hipGraph captures, tight serving loops, harnesses that
// forbid hipMalloc) use this file-scope array instead.  Every GPU the library
// runs on holds one copy, NOALLOC_MAX_N * 76 bytes (19 MiB), until the process
// ends.  Calls on one GPU share it: two downloads that overlap on one GPU
// (from two contexts or threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gdem_kernels.h"

namespace gdem {

// x, v, w: 3 doubles each per particle, then the activity flags as ints
static __device__ __align__(256) double scratch[(size_t)NOALLOC_MAX_N * 9 + NOALLOC_MAX_N / 2];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
}  // namespace

// Resolves the array's device address on the current device (once per
// device).  The host-side name `scratch` is not a device address; only the
// pointer kept here is passed to kernels and copies.
hipError_t noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[device] = static_cast<double*>(p);
  return hipSuccess;
}

NoallocStage noalloc_stage(int device) {
  NoallocStage s{};
  if (device < 0 || device >= kMaxDevices) return s;
  std::lock_guard<std::mutex> lk(g_mu);
  double* base = g_ptr[device];
  if (!base) return s;
  s.x = base;
  s.v = base + (size_t)NOALLOC_MAX_N * 3;
  s.w = base + (size_t)NOALLOC_MAX_N * 6;
  s.act = reinterpret_cast<int*>(base + (size_t)NOALLOC_MAX_N * 9);
  return s;
}

}  // namespace gdem
