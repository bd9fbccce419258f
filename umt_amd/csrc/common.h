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

// Zone-blocked copy of a per-corner array of k components (hex zones of 8
// corners, corner-major as Teton stores it): component j of corner 8 z + c
// goes to (z k + j) 8 + c, so each zone's block is contiguous and holds
// each component for its eight corners side by side.
template <typename T>
std::vector<T> zone_blocked(const T* in, int k, size_t ncorner) {
  std::vector<T> out((size_t)k * ncorner);
  for (size_t cc = 0; cc < ncorner; ++cc) {
    const size_t z = cc / 8, c = cc % 8;
    for (int j = 0; j < k; ++j) out[(z * k + j) * 8 + c] = in[cc * k + j];
  }
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
