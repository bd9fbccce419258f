// Device arithmetic helpers shared by the sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace umt {

// Quotient a / b. FP32 multiplies by the hardware reciprocal estimate
// (v_rcp_f32, 1 ulp) instead of running the correctly rounded divide
// sequence of about ten instructions; the FP32 sweep's error budget is far
// above that. FP64 refines v_rcp_f64 by one Newton-Raphson step
// r <- r + r (1 - b r), forms q = a r and corrects it once with the exact
// residual, q + r (a - b q) (Markstein's final step). v_rcp_f64's relative
// error is about 2^-24 (measured on MI355X: at most 2^-24.4 over 2^24 random
// operands), one step leaves about 2^-48, and the correction then rounds
// correctly but for quotients within about 2^-96 of a rounding boundary: on
// 2^24 random pairs with exponents in [-200, 200] the result equalled the
// IEEE quotient a / b in every case. That is the arithmetic of the
// compiler's divide without its second Newton step, operand scaling
// (v_div_scale) and special-value fixup (v_div_fixup), for finite operands
// away from the ends of the exponent range, as here: 6 instructions
// instead of 11. (Multiplying by the reciprocal without the correction adds
// a rounding: a case with cancelling corner terms then moved the FP64 sweep
// by 1.02e-12 (elementwise) from the oracle, past its 1e-12 bound.)
template <typename Real>
__device__ __forceinline__ Real quot(Real a, Real b);
template <>
__device__ __forceinline__ float quot<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
template <>
__device__ __forceinline__ double quot<double>(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(r, fma(-b, r, 1.0), r);
  const double q = a * r;
  return fma(r, fma(-b, q, a), q);
}

// Max that keeps a NaN once one is seen. The convergence kernels reduce
// non-negative doubles through atomicMax on their bit patterns; a positive
// quiet NaN's pattern is above +Inf's and every finite value's, so a NaN
// that reaches the atomic wins and the host's finiteness check sees it.
__device__ __forceinline__ double nan_keep_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return fmax(a, b);
}
constexpr unsigned long long kNaNBits = 0x7ff8000000000000ull;

}  // namespace umt
