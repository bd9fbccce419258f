// Internal helpers shared by the library's host translation units: error
// reporting (the message umt_last_error returns), device allocation and the
// HIP status check.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace umt {

// Records the calling thread's last error message and returns code.
int fail(int code, const char* fmt, ...);

// 0 when every zone has the canonical hex corner topology the kernels use.
int check_hex_topology(int Z, int C, const int* numCorner, const int* cOffSet, const int* nCFaces,
                       const int* cEZ);

template <typename T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) return 0;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess)
    return fail(100 + (int)e, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  return 0;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace umt

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return umt::fail(100 + (int)e_, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                              \
  } while (0)
