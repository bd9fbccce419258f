// Internal helpers shared by the library's host translation units: error
// reporting (the message umt_last_error returns), device allocation and the
// HIP status check.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

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

// Component-major copy of an array of n records of k components each:
// out[j * n + i] = in[i * k + j] (Teton's corner arrays are record-major).
template <typename T>
std::vector<T> component_major(const T* in, int k, size_t n) {
  std::vector<T> out((size_t)k * n);
  for (size_t i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) out[(size_t)j * n + i] = in[i * k + j];
  return out;
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
