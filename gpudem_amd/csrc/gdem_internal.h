// Host-side definitions shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gdem.h"

struct gdem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace gdem {

inline int fail(gdem_ctx* c, int code, const std::string& what) {
  if (c) c->err = what;
  return code;
}

#define HIPCHK(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::gdem::fail((ctx), e_ == hipErrorOutOfMemory ? GDEM_E_OOM : GDEM_E_HIP,   \
                          std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

// Owns a set of device allocations.
struct DevArena {
  std::vector<void*> ptrs;
  DevArena() = default;
  DevArena(const DevArena&) = delete;
  DevArena& operator=(const DevArena&) = delete;
  ~DevArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T** out, size_t n) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
      ptrs.push_back(p);
      *out = static_cast<T*>(p);
    }
    return e;
  }
};

}  // namespace gdem
