// Device arithmetic helpers shared by the sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace umt {

// Quotient a / b. FP32 multiplies by the hardware reciprocal estimate
// (v_rcp_f32, 1 ulp) instead of running the correctly rounded divide
// sequence of about ten instructions; the FP32 sweep's error budget is far
// above that. FP64 divides correctly rounded: multiplying by a reciprocal
// adds a rounding, and a case with cancelling corner terms then moved the
// FP64 sweep by 1.02e-12 (elementwise) from the oracle, at its 1e-12 bound.
template <typename Real>
__device__ __forceinline__ Real quot(Real a, Real b);
template <>
__device__ __forceinline__ float quot<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
template <>
__device__ __forceinline__ double quot<double>(double a, double b) {
  return a / b;
}

}  // namespace umt
