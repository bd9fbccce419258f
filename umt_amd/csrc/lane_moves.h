// Lane moves and group vectors shared by the sweep kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace umt {

// DPP moves of 32-bit lanes (64-bit values move as two halves). kCtrl is
// the dpp_ctrl encoding: quad_perm patterns move values inside groups of
// four lanes, row shifts and mirrors within a row of sixteen. A source
// outside the row reads 0 (bound_ctrl), so no old value has to be set up.
// The sweep kernels call these with every lane of the wave active: a lane
// whose source lane is switched off reads 0, so a move must not sit in a
// branch, nor be an operand of ?:, && or || (evaluated on some lanes only).
template <int kCtrl>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, true);
}
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_i32<kCtrl>((int)b), hi = dpp_i32<kCtrl>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int kCtrl>
__device__ __forceinline__ float dpp_move(float x) {
  return __int_as_float(dpp_i32<kCtrl>(__float_as_int(x)));
}
template <int kCtrl>
__device__ __forceinline__ int dpp_move(int x) {
  return dpp_i32<kCtrl>(x);
}
constexpr int kQuadPrev = 0x93;  // quad_perm [3,0,1,2]: value of lane (i - 1) mod 4
constexpr int kQuadNext = 0x39;  // quad_perm [1,2,3,0]: value of lane (i + 1) mod 4
constexpr int kRowShl4 = 0x104;
constexpr int kRowShr4 = 0x114;
constexpr int kRowHalfMirror = 0x141;  // row_half_mirror: lane i <-> 7 - i within each 8 lanes

// K consecutive groups of one corner as one aligned vector access.
template <typename Real, int K>
struct GroupVec {
  Real v[K];
};
template <typename Real, int K>
__device__ __forceinline__ GroupVec<Real, K> load_groups(const Real* p) {
  GroupVec<Real, K> r;
  __builtin_memcpy(&r, __builtin_assume_aligned(p, K * sizeof(Real)), sizeof r);
  return r;
}
template <typename Real, int K>
__device__ __forceinline__ void store_groups(Real* p, const GroupVec<Real, K>& r) {
  __builtin_memcpy(__builtin_assume_aligned(p, K * sizeof(Real)), &r, sizeof r);
}

}  // namespace umt
