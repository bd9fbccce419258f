// CDNA4 kernels for the upstream-corner-balance (UCB) xyz sweep.
//
// One lane owns one corner of one (zone, angle, group) triple: eight
// consecutive lanes are the eight corners of a hex zone for one direction and
// group, and a wave holds eight such (item, group) pairs, groups of one item
// next to each other, so the loads of a corner's group vector (S, psi: groups
// fastest) are contiguous across the pairs. The corners of a zone exchange
// values with DPP lane moves inside their 8-lane group. Work is handed out
// per hyperplane level: a level holds every (zone, angle) pair whose upstream
// zones lie in earlier levels, for all angles at once, so one launch per
// level keeps the whole GPU busy.
//
// The corner arithmetic follows snac/SweepUCBxyz.F90:119-306: volume terms,
// incoming FP-face fluxes, the EZ coupling term sez with fouralpha = 1.82,
// and the per-corner solve psi_c = src_c / (sum of outgoing areas + sigma V_c).
// Of a zone's 12 EZ edges, each of the 8 on the two quad rings has its term
// computed once, by the lane that owns it (ring_term); each of the 4
// vertical edges by its upwind corner, as the reference does (sez_term).
// The within-zone corner solve does not need the nextC order: corners are
// relaxed in a fixed order, each pulling the EZ couplings of its upstream
// neighbours, until every corner has been computed from final upstream
// values (the result of a triangular solve does not depend on which
// topological order produced it).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "device_math.h"
#include "kernels.h"
#include "lane_moves.h"

namespace umt {

// Zeros read in place of the upwind psi of a face that brings none (at
// least one group vector of the widest groups-per-lane variant). Not const: a
// const array sits in the constant address space, and a select between it
// and a psi address makes the upwind loads flat loads, whose waits drain the
// scalar and LDS counter too; in the global address space they stay global
// loads. Nothing writes it.
constexpr int kMaxGroupsPerLane = 8;
// 64 values: the zone layout reads element `lane` of it (lane < 64).
__device__ __align__(64) double kZeroGroups[64] = {};

// The kernels assume the canonical hex topology of the tiled-mesh generator:
// the EZ neighbour of local corner c (0-based) in corner-face slots 0..2 is
//   {3,1,4} {0,2,5} {1,3,6} {2,0,7} {5,7,0} {6,4,1} {7,5,2} {4,6,3}
// for c = 0..7 (corner-face chaining of aux/setTetonZone.F90:172-189). The
// runtime checks every zone against this table before launching them.
// Omega . A for corner-face vector components stored stride apart (the
// device keeps geometry zone-blocked: component k of corner 8 z + c at
// (9 z + k) 8 + c).
template <typename Real>
__device__ __forceinline__ Real dot3(Real ox, Real oy, Real oz, const Real* a, long stride) {
  return ox * a[0] + oy * a[stride] + oz * a[2 * stride];
}

// Cross-lane moves inside the 8-lane group of one (item, group) pair. Lane
// l holds corner l for l < 4 and corner 11 - l for l >= 4 (the top quad in
// reverse), so every EZ neighbour is one uniform DPP move (lane_moves.h):
// slot 0 (kEZ[c][0]) is the previous lane of the quad (quad_perm
// [3,0,1,2]), slot 1 the next lane (quad_perm [1,2,3,0]), slot 2 (c ^ 4)
// the mirror lane of the half row (row_half_mirror: lane l <-> 7 - l).
__device__ __forceinline__ int corner_of_lane(int l) { return l < 4 ? l : 11 - l; }

// Value held by this lane's EZ neighbour in corner-face slot F.
template <int F, typename T>
__device__ __forceinline__ T from_ez(T x) {
  if constexpr (F == 0) return dpp_move<kQuadPrev>(x);
  else if constexpr (F == 1) return dpp_move<kQuadNext>(x);
  else return dpp_move<kRowHalfMirror>(x);
}

// The EZ coupling term of EZ face f of a corner (SweepUCBxyz.F90:217-252),
// used for the vertical edges (slot 2), fouralpha = 1.82, t = sigma v:
//   gnum = A^2 (1.82 t^2 + A (4 t + 3 A)), gden = v (4 t^3 + A (6 t^2 + 2 A (2 t + A))),
//   sez = (v gnum (sigma psiOpp - Qu) + A/2 gden (Qu - Qd)) / (gnum + gden sigma)
// when the face is outgoing (A = Omega.A_ez > 0) and the opposite FP face
// incoming, A (Qu - Qd) / (2 sigma) when the opposite face is not incoming,
// and 0 on a face that is not outgoing.
//
// FP64 evaluates all three cases with the first expression and no selects:
// A = max(Omega.A_ez, 0) and gnum scaled by mOpp (1 with an incoming
// opposite face, else 0). gnum = 0 leaves A/2 gden (Qu - Qd) / (gden sigma),
// the second case; A = 0 leaves 0 / (4 v t^3 sigma) = 0, the third (t^3
// stays far above the FP64 underflow). The first case is bit for bit the
// reference expression. FP32 keeps the selects (an FP32 select is one
// instruction, and 4 v t^3 sigma can underflow in FP32).
template <typename Real>
__device__ __forceinline__ Real sez_term(Real aez, Real mOpp, Real t, Real t2, Real sig, Real v,
                                         Real Qu, Real Qd, Real psiOpp) {
  if constexpr (sizeof(Real) == 8) {
    const Real A = aez > 0 ? aez : (Real)0;
    const Real gnum = mOpp * (A * A) * ((Real)1.82 * t2 + A * ((Real)4 * t + (Real)3 * A));
    const Real gden = v * ((Real)4 * t * t2 + A * ((Real)6 * t2 + (Real)2 * A * ((Real)2 * t + A)));
    return quot<Real>(v * gnum * (sig * psiOpp - Qu) + (Real)0.5 * A * gden * (Qu - Qd),
                      gnum + gden * sig);
  } else {
    const Real A = aez;
    const Real A2 = A * A;
    const Real gnum = A2 * ((Real)1.82 * t2 + A * ((Real)4 * t + (Real)3 * A));
    const Real gden = v * ((Real)4 * t * t2 + A * ((Real)6 * t2 + (Real)2 * A * ((Real)2 * t + A)));
    const Real withOpp = quot<Real>(v * gnum * (sig * psiOpp - Qu) + (Real)0.5 * A * gden * (Qu - Qd),
                                    gnum + gden * sig);
    const Real noOpp = A * quot<Real>((Real)0.5, sig) * (Qu - Qd);
    const Real x = mOpp > 0 ? withOpp : noOpp;
    return aez > 0 ? x : (Real)0;
  }
}

// The EZ coupling term of the ring edge this lane owns: the EZ face shared
// with its slot-1 neighbour (the next lane of the quad, which sees the face
// in slot 0). The face areas of the two corners are equal and opposite
// (geometryUCBxyz.F90:166-178), so A = |aez| is the upwind corner's area
// whichever corner that is. vU, tU = sigma vU, t2U, mU and wU = sigma
// psiOpp - Qu belong to the upwind corner; dQ = Q(this) - Q(next) and the
// sign of aez (this corner's area) carry the orientation:
// 0.5 A (Qu - Qd) = 0.5 aez dQ exactly. The value is added to the upwind
// corner's source and taken off the downwind one's; a face parallel to the
// direction (aez = 0) gives 0. Computing each ring edge once, in the lane
// that owns it, replaces the two evaluations (one per corner, one of them
// 0) of the per-corner form sez_term, whose select-free FP64 arithmetic the
// first expression here follows.
template <typename Real>
__device__ __forceinline__ Real ring_term(Real aez, Real mU, Real tU, Real t2U, Real sig, Real vU,
                                          Real wU, Real dQ) {
  const Real A = fabs(aez), halfA = (Real)0.5 * aez;
  const Real poly = (Real)1.82 * t2U + A * ((Real)4 * tU + (Real)3 * A);
  const Real gden = vU * ((Real)4 * tU * t2U + A * ((Real)6 * t2U + (Real)2 * A * ((Real)2 * tU + A)));
  if constexpr (sizeof(Real) == 8) {
    const Real gnum = mU * (A * A) * poly;
    return quot<Real>(vU * gnum * wU + halfA * gden * dQ, gnum + gden * sig);
  } else {
    const Real gnum = (A * A) * poly;
    const Real withOpp = quot<Real>(vU * gnum * wU + halfA * gden * dQ, gnum + gden * sig);
    const Real noOpp = aez * quot<Real>((Real)0.5, sig) * dQ;
    const Real x = mU > 0 ? withOpp : noOpp;
    return A > 0 ? x : (Real)0;
  }
}

// One lane: one corner c of one zone for one direction and K consecutive
// groups. Geometry, dot products and the zone's EZ orientation are shared by
// the K groups; everything group-dependent is a K-vector in registers.
// Item `item` of the level, groups g0 .. g0+K-1, this lane's corner. Lanes
// with valid == false still take part in the cross-lane moves and the wave
// vote (they work on a valid item) but store nothing.
template <typename Real, int K, bool kSingleAngle>
__device__ __forceinline__ void sweep_corner(const SweepLevelArgs<Real>& p, int item, int g0,
                                             bool valid) {
  static_assert(K <= kMaxGroupsPerLane, "the zero group vector is shorter than K");
  const int c = corner_of_lane(threadIdx.x & 7);
  const int G = p.G;
  const int2 it = p.items[item];
  // A zone whose corners depend on each other in a cycle comes as ~zone.
  const bool cyclic = it.x < 0;
  const int z = cyclic ? ~it.x : it.x, a = it.y;
  const long cc = 8L * z + c;
  const Real ox = (Real)p.omega[3 * a + 0];
  const Real oy = (Real)p.omega[3 * a + 1];
  const Real oz = (Real)p.omega[3 * a + 2];
  const long angleBase = kSingleAngle ? 0L : (long)(a - p.angle0) * p.psiAngleStride;
  Real* __restrict__ psiA = p.psi + angleBase;

  const Real v = p.vol[cc];
  // Geometry is zone-blocked on the device (component k = 3 face + dim of
  // corner c at (9 z + k) 8 + c, cFP face f at (3 z + f) 8 + c): each load
  // reads the same component of the zone's eight corners side by side, and
  // a zone's block is contiguous.
  const Real* Afz = p.Afp + 72L * z + c;
  const Real* Aez = p.Aez + 72L * z + c;
  Real afp[3], aez[3];
  int cfp[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    afp[f] = dot3(ox, oy, oz, Afz + 24 * f, 8);
    aez[f] = dot3(ox, oy, oz, Aez + 24 * f, 8);
    cfp[f] = p.cFP[24L * z + 8 * f + c] - 1;
  }
  if (p.parTol > 0) {
    // Directions lying in a mesh face leave rounding noise of either sign in
    // Omega.A; such faces are parallel here as in the schedule, so nothing is
    // read across them and they cannot select the opposite-face EZ branch.
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const Real* af = Afz + 24 * f;
      const Real* ae = Aez + 24 * f;
      if (fabs(afp[f]) <= p.parTol * (fabs(af[0]) + fabs(af[8]) + fabs(af[16]))) afp[f] = 0;
      if (fabs(aez[f]) <= p.parTol * (fabs(ae[0]) + fabs(ae[8]) + fabs(ae[16]))) aez[f] = 0;
    }
  }
  Real outArea = 0;
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (afp[f] > 0) outArea += afp[f];
    if (aez[f] > 0) outArea += aez[f];
  }
  const GroupVec<Real, K> sigv_ = load_groups<Real, K>(p.sigt + (long)z * G + g0);
  GroupVec<Real, K> Q = load_groups<Real, K>(p.S + cc * G + g0);
  if (p.tauPsi) {
    // Time-dependent source tau Psi(g, c, Angle) (SweepUCBxyz.F90:121): the
    // drop-in passes this angle's Psi (G, C); the all-angle sweep reads the
    // previous cycle's psi of angle a, stored like psi with every angle.
    const Real* tpA = kSingleAngle ? p.tauPsi : p.tauPsi + (long)a * p.psiAngleStride;
    const GroupVec<Real, K> tp = load_groups<Real, K>(tpA + cc * G + g0);
#pragma unroll
    for (int k = 0; k < K; ++k) Q.v[k] += p.tau * tp.v[k];
  }
  // Incoming FP faces (SweepUCBxyz.F90:139-161). An incoming boundary face
  // reads the boundary tail of psi when the sweep has incident fluxes (the
  // single-angle drop-in, or an all-angle sweep whose tails were loaded
  // before its levels); otherwise it brings zero (vacuum) and the tail is
  // only written. noInflow sweeps every zone on its own (no upwind reads).
  GroupVec<Real, K> psf[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const bool rd = afp[f] < 0 && (kSingleAngle || p.boundaryIn || cfp[f] < p.C) && !p.noInflow;
    if constexpr (sizeof(Real) == 8) {
      // FP64: every face loads, a face without upwind psi from a zero
      // buffer (no branches, and fewer instructions; FP32 keeps the branch,
      // which needs fewer registers there).
      psf[f] = load_groups<Real, K>(rd ? psiA + (long)cfp[f] * G + g0 : (const Real*)kZeroGroups);
    } else if (rd) {
      psf[f] = load_groups<Real, K>(psiA + (long)cfp[f] * G + g0);
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) psf[f].v[k] = 0;
    }
  }
  // Per group: volume and incoming FP terms (SweepUCBxyz.F90:119-161), then
  // the EZ coupling terms computed by the upwind corner (:182-252), pulled
  // with the opposite sign by the downwind one.
  // Incoming FP areas, +0 on the other faces (whose psf is 0, so their term
  // is s - (+0) = s, as the reference skips them).
  Real afpIn[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) afpIn[f] = afp[f] < 0 ? afp[f] : (Real)0;
  // 1 where the FP face opposite EZ face f is incoming, else 0.
  Real mOpp[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) mOpp[f] = afp[(f + 1) % 3] < 0 ? (Real)1 : (Real)0;
  // The ring edge this lane owns (see ring_term): the upwind corner's
  // volume and opposite-face flag, and the signs with which this lane takes
  // its own ring edge (slot 1) and the previous lane's (its slot 0).
  // The moves run on every lane before the selects: as an operand of ?:
  // a move would run only on the lanes taking that branch, and a lane whose
  // source lane is switched off reads 0.
  const bool up1 = aez[1] > 0;
  const Real vN = from_ez<1>(v), mN = from_ez<1>(mOpp[0]);
  const Real vU = up1 ? v : vN;
  const Real mU = up1 ? mOpp[1] : mN;
  const Real sg1 = up1 ? (Real)1 : (Real)-1, sg0 = aez[0] > 0 ? (Real)1 : (Real)-1;
  Real src[K], rden[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const Real sig = sigv_.v[k];
    const Real q = Q.v[k];
    // Volume term and incoming FP terms (SweepUCBxyz.F90:139-161): each
    // face product rounded on its own, the first taken off V q in one fused
    // step. Both forms give the same bits; the masked one saves the FP64
    // kernel 18 selects (1% faster), the FP32 kernel is 1% faster with the
    // selects.
    Real s;
    if constexpr (sizeof(Real) == 8) {
#pragma clang fp contract(off)
      s = fma(v, q, -(afpIn[0] * psf[0].v[k]));
      s = s - afpIn[1] * psf[1].v[k];
      s = s - afpIn[2] * psf[2].v[k];
    } else {
      s = v * q;
#pragma unroll
      for (int f = 0; f < 3; ++f) s = s - (afp[f] < 0 ? afp[f] * psf[f].v[k] : (Real)0);
    }
    rden[k] = quot<Real>(1, outArea + sig * v);
    const Real t = sig * v, t2 = t * t;
    // Ring edge: sigma psiOpp - Q of this corner on its slot-1 face
    // (opposite FP face 2) and of the next corner on its slot-0 face
    // (opposite FP face 1, computed there and moved here).
    const Real w1 = sig * psf[2].v[k] - q;
    const Real wN = from_ez<1>(sig * psf[1].v[k] - q);
    const Real tU = sig * vU;
    const Real e = ring_term<Real>(aez[1], mU, tU, tU * tU, sig, vU, up1 ? w1 : wN,
                                   q - from_ez<1>(q));
    // Vertical edge (slot 2, the mirror lane): each corner computes it when
    // upwind and the mirror lane takes it off.
    const Real s2 = sez_term<Real>(aez[2], mOpp[2], t, t2, sig, v, q, from_ez<2>(q), psf[0].v[k]);
    s = fma(sg1, e, s);
    s = fma(sg0, from_ez<0>(e), s);
    s += s2;
    s -= from_ez<2>(s2);
    src[k] = s;
  }

  // Corner solve psi_c = (src_c + sum over upstream EZ neighbours of
  // |aez| psi_n) / den_c, relaxed until every corner saw final upstream values.
  Real k0 = aez[0] < 0 ? -aez[0] : (Real)0;
  Real k1 = aez[1] < 0 ? -aez[1] : (Real)0;
  Real k2 = aez[2] < 0 ? -aez[2] : (Real)0;
  GroupVec<Real, K> psi;
#pragma unroll
  for (int k = 0; k < K; ++k) psi.v[k] = src[k] * rden[k];
  if (__any(cyclic)) {
    // A cyclic zone is solved once from the previous values of its own
    // corners (SweepUCBxyz.F90:281-295), which psi holds until the zone is
    // computed: its corners are on the angle's cycle list. The 8 lanes of
    // an item are all cyclic or all not, so the moves stay inside the item.
    // Its lanes then take no upstream terms, a unit reciprocal and the
    // solve as source, so the relaxation below reproduces it exactly.
    GroupVec<Real, K> old{};
    if (cyclic) old = load_groups<Real, K>(psiA + cc * G + g0);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Real x = cyclic ? old.v[k] : (Real)0;
      const Real n0 = from_ez<0>(x), n1 = from_ez<1>(x),
                 n2 = from_ez<2>(x);
      const Real once = (src[k] + k0 * n0 + k1 * n1 + k2 * n2) * rden[k];
      psi.v[k] = cyclic ? once : psi.v[k];
      src[k] = cyclic ? once : src[k];
      rden[k] = cyclic ? (Real)1 : rden[k];
    }
    if (cyclic) k0 = k1 = k2 = 0;
  }
  int fin = (k0 == 0 && k1 == 0 && k2 == 0) ? 1 : 0;
  for (int iter = 0; iter < 8; ++iter) {
    if (__all(fin)) break;
    const int f0 = from_ez<0>(fin), f1 = from_ez<1>(fin),
              f2 = from_ez<2>(fin);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Real x = psi.v[k];
      const Real n0 = from_ez<0>(x), n1 = from_ez<1>(x),
                 n2 = from_ez<2>(x);
      psi.v[k] = (src[k] + k0 * n0 + k1 * n1 + k2 * n2) * rden[k];
    }
    fin = (k0 == 0 || f0) && (k1 == 0 || f1) && (k2 == 0 || f2);
  }

  if (valid) {
    store_groups<Real, K>(psiA + cc * G + g0, psi);
    if (kSingleAngle)
#pragma unroll
      for (int k = 0; k < K; ++k) p.phi[cc * G + g0 + k] += p.quadwt * (double)psi.v[k];
#pragma unroll
    for (int f = 0; f < 3; ++f)
      if (afp[f] > 0 && cfp[f] >= p.C) store_groups<Real, K>(psiA + (long)cfp[f] * G + g0, psi);
  }
}

template <typename Real, int K, bool kSingleAngle>
__global__ __launch_bounds__(256) void ucb_sweep_level(SweepLevelArgs<Real> p) {
  const unsigned pair = (blockIdx.x * blockDim.x + threadIdx.x) >> 3;
  const unsigned blocksPerItem = (unsigned)(p.G / K);
  int item = (int)fast_div(pair, p.perItem);
  const int g0 = K * (int)(pair - (unsigned)item * blocksPerItem);
  const bool valid = item < p.nItems;
  if (!valid) item = 0;
  sweep_corner<Real, K, kSingleAngle>(p, item, g0, valid);
}

// ---------------------------------------------------------------------------
// Zone layout: one lane owns all eight corners of one (zone, angle) item for
// one group, and W consecutive lanes (a slot: W = 16, 32 or 64, dividing G)
// hold W consecutive groups of the same item, 64 / W slots per wave. Every
// corner-to-corner coupling of the zone is then inside one lane's registers:
// no lane moves, each of the zone's 12 EZ edge terms evaluated once (the 4
// vertical ones only at their upwind corner, behind a branch that is uniform
// over the slot), and the corner solve as in-register Gauss-Seidel passes in
// corner order until every corner was computed from final upstream values.
// Every quantity is the corner layout's expression in the same order (the
// two agree to a few ulps: the compiler fuses multiply-adds differently in
// places). Everything that depends on the item but not on the group
// (Omega.A of the 48 corner faces, the signs, coefficients and upwind
// operands derived from them, volumes, upwind psi offsets, the number of
// Gauss-Seidel passes) is computed once per slot by its first 8 lanes, one
// corner each, and read by all W lanes from LDS (one address per slot:
// broadcast). All-angle sweeps only (the per-angle drop-in keeps the corner
// layout).

// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>): a loop
// whose index is a compile-time constant in its body from the start, so
// register arrays indexed by it (and by constant functions of it) are split
// into registers before any other optimisation sees them.
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// EZ neighbour of corner c in corner-face slot f, the canonical hex table
// {3,1,4} {0,2,5} {1,3,6} {2,0,7} {5,7,0} {6,4,1} {7,5,2} {4,6,3} as
// arithmetic, so that with c and f known at compile time it folds to a
// constant (register arrays indexed by it stay in registers).
__device__ __forceinline__ constexpr int kEZ(int c, int f) {
  return f == 2 ? (c ^ 4)
                : (c < 4 ? ((c + (f == 0 ? 3 : 1)) & 3) : 4 + ((c + (f == 0 ? 1 : 3)) & 3));
}

// A compiler-only memory fence: loads after it are not merged with loads
// before it, so values in LDS are read again where they are used instead of
// being held in registers in between.
__device__ __forceinline__ void reread_lds() { asm volatile("" ::: "memory"); }
// Keeps the scheduler from interleaving the arithmetic of consecutive
// corners or edges: interleaved, their temporaries are live together and
// take more registers than the latency hiding is worth.
__device__ __forceinline__ void one_at_a_time() { __builtin_amdgcn_sched_barrier(0); }
// x, computed here and opaque to the compiler from here on: it cannot sink
// the computation of x to a later use (which would keep its operands live
// across the phases in between), nor turn a select of two such values read
// from LDS into a read through a selected address.
template <typename T>
__device__ __forceinline__ T pinned(T x) {
  asm volatile("" : "+v"(x));
  return x;
}

// One slot's item as its lanes read it: every quantity that depends on the
// item but not on the group, computed once by the slot's corner lanes in the
// form the group lanes use it.
template <typename Real>
struct ZoneRec {
  Real afpIn[8][3];   // Omega.A_fp of incoming FP faces, 0 on the others
  Real kup[8][3];     // |Omega.A_ez| of upstream EZ faces, 0 on the others (all 0
                      // for a cyclic zone, which takes no Gauss-Seidel passes)
  Real kOnce[8][3];   // the same for a cyclic zone's one solve from lagged values
  Real vol[8];
  Real outArea[8];    // sum of the outgoing FP and EZ areas of each corner
  Real ringA[8];      // ring edge owned by c (to its slot-1 neighbour): c's Omega.A_ez
  Real ringV[8];      //   the upwind corner's volume
  Real ringM[8];      //   1 where the upwind corner's FP face opposite the edge is incoming
  Real sg0[8];        // +1 / -1: c upwind / not on its slot-0 EZ face
  Real sg1[8];        // the same on its slot-1 EZ face
  Real vertA[4];      // vertical edge (i, i + 4): the upwind corner's Omega.A_ez (<= 0: none)
  Real vertV[4];      //   its volume
  Real vertM[4];      //   1 where its FP face 0 (opposite the edge) is incoming
  Real vertS[8];      // +1 where c adds the vertical edge's term, -1 where it takes it off
  long off[8][3];     // element offset from psiA + lane of the value each FP face brings
                      // in: the upwind psi of an incoming face that reads one, else a
                      // zero of kZeroGroups
  int cfp[8][3];      // corner or boundary element across each FP face, 0-based
  int flags[8];       // bit 0: c upwind on its own ring edge; bit 1: the bottom corner of
                      // vertical edge c (c < 4) is its upwind one; bits 2 + f: psi of c
                      // goes to the boundary element across outgoing FP face f
};

template <typename Real, int W>
__global__ __launch_bounds__(256) void ucb_sweep_level_zone(SweepLevelArgs<Real> p) {
  static_assert(W == 16 || W == 32 || W == 64, "a slot is 16, 32 or 64 lanes");
  constexpr int kSlots = 256 / W;
  __shared__ ZoneRec<Real> rec[kSlots];
  const int lane = threadIdx.x % W;  // group within the slot's block of W groups
  ZoneRec<Real>& R = rec[threadIdx.x / W];
  const unsigned task = (blockIdx.x * 256u + threadIdx.x) / W;
  const unsigned blocksPerItem = (unsigned)(p.G / W);
  int item = (int)fast_div(task, p.perItem);
  const int g = W * (int)(task - (unsigned)item * blocksPerItem) + lane;
  // Lanes past the level's last item work on item 0 and store nothing.
  const bool valid = item < p.nItems;
  if (!valid) item = 0;
  const int2 it = p.items[item];
  // A zone whose corners depend on each other in a cycle comes as ~zone.
  const bool cyclic = it.x < 0;
  const int z = cyclic ? ~it.x : it.x, a = it.y;
  const int G = p.G;
  Real* __restrict__ psiA = p.psi + (long)(a - p.angle0) * p.psiAngleStride;

  // The group's cross section and source need only the item: their loads
  // are in flight while the prologue runs.
  const Real sig = p.sigt[(long)z * G + g];
  Real Q[8];
  static_for<8>([&](auto C) {
    constexpr int c = C;
    Q[c] = p.S[(8L * z + c) * G + g];
  });
  if (p.tauPsi) {
    // Time-dependent source tau Psi (SweepUCBxyz.F90:121): the previous
    // cycle's psi of angle a, stored like psi with every angle.
    const Real* tpA = p.tauPsi + (long)a * p.psiAngleStride;
    static_for<8>([&](auto C) {
      constexpr int c = C;
      Q[c] += p.tau * tpA[(8L * z + c) * G + g];
    });
  }

  // Prologue: lane l of a slot computes corner c = l mod 8 (the corner
  // layout's per-lane geometry; lanes 8 and up repeat the first eight, which
  // keeps every lane active for the lane reads), and lanes 0..7 store it.
  const int c = lane & 7;
  const int slotBase = (int)(threadIdx.x % 64) - lane;  // first lane of the slot in its wave
  unsigned need = 0;  // upstream EZ neighbours of corner c (lanes 0..7)
  {
    const Real ox = (Real)p.omega[3 * a + 0];
    const Real oy = (Real)p.omega[3 * a + 1];
    const Real oz = (Real)p.omega[3 * a + 2];
    const Real* Afz = p.Afp + 72L * z + c;
    const Real* Aez = p.Aez + 72L * z + c;
    Real afp[3], aez[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      afp[f] = dot3(ox, oy, oz, Afz + 24 * f, 8);
      aez[f] = dot3(ox, oy, oz, Aez + 24 * f, 8);
    }
    if (p.parTol > 0) {
      // Directions lying in a mesh face leave rounding noise of either sign
      // in Omega.A; such faces are parallel here as in the schedule.
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        const Real* af = Afz + 24 * f;
        const Real* ae = Aez + 24 * f;
        if (fabs(afp[f]) <= p.parTol * (fabs(af[0]) + fabs(af[8]) + fabs(af[16]))) afp[f] = 0;
        if (fabs(aez[f]) <= p.parTol * (fabs(ae[0]) + fabs(ae[8]) + fabs(ae[16]))) aez[f] = 0;
      }
    }
    Real outArea = 0;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (afp[f] > 0) outArea += afp[f];
      if (aez[f] > 0) outArea += aez[f];
    }
    const Real v = p.vol[8L * z + c];
    int cf[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) cf[f] = p.cFP[24L * z + 8 * f + c] - 1;
    // Operands from the other corner of this corner's ring edge (slot-1
    // neighbour) and vertical edge (c ^ 4), read from their lanes.
    const int n1 = c < 4 ? (c + 1) & 3 : 4 + ((c + 3) & 3);
    const Real vN = __shfl(v, slotBase + n1, 64);
    const bool inN1 = __shfl(afp[1] < 0 ? 1 : 0, slotBase + n1, 64) != 0;
    const Real vV = __shfl(v, slotBase + (c ^ 4), 64);
    const Real aV = __shfl(aez[2], slotBase + (c ^ 4), 64);
    const bool inV0 = __shfl(afp[0] < 0 ? 1 : 0, slotBase + (c ^ 4), 64) != 0;
    const long zeroOff = ((long)(uintptr_t)kZeroGroups - (long)(uintptr_t)psiA) / (long)sizeof(Real);
    const int g0 = g - lane;  // first group of the slot's block
    if (lane < 8) {
      int flags = 0;
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        R.afpIn[c][f] = afp[f] < 0 ? afp[f] : (Real)0;
        const Real k = aez[f] < 0 ? -aez[f] : (Real)0;
        R.kOnce[c][f] = k;
        R.kup[c][f] = cyclic ? (Real)0 : k;
        R.cfp[c][f] = cf[f];
        // The corner layout's read rule for an incoming FP face.
        const bool rd = afp[f] < 0 && (p.boundaryIn || cf[f] < p.C) && !p.noInflow;
        R.off[c][f] = rd ? (long)cf[f] * G + g0 : zeroOff;
        if (afp[f] > 0 && cf[f] >= p.C) flags |= 4 << f;
        if (aez[f] < 0 && !cyclic) need |= 1u << (f == 2 ? (c ^ 4) : (f == 1 ? n1 : (c < 4 ? (c + 3) & 3 : 4 + ((c + 1) & 3))));
      }
      R.vol[c] = v;
      R.outArea[c] = outArea;
      const bool up1 = aez[1] > 0;
      R.ringA[c] = aez[1];
      R.ringV[c] = up1 ? v : vN;
      R.ringM[c] = (up1 ? afp[2] < 0 : inN1) ? (Real)1 : (Real)0;
      if (up1) flags |= 1;
      R.sg0[c] = aez[0] > 0 ? (Real)1 : (Real)-1;
      R.sg1[c] = up1 ? (Real)1 : (Real)-1;
      // Vertical edge (i, i + 4); Omega.A_ez of i + 4 is minus that of i.
      const bool upBottom = c < 4 ? aez[2] > 0 : aez[2] < 0;
      R.vertS[c] = (c < 4) == upBottom ? (Real)1 : (Real)-1;
      if (c < 4) {
        R.vertA[c] = upBottom ? aez[2] : aV;
        R.vertV[c] = upBottom ? v : vV;
        R.vertM[c] = (upBottom ? afp[0] < 0 : inV0) ? (Real)1 : (Real)0;
        if (upBottom) flags |= 2;
      }
      R.flags[c] = flags;
    }
  }
  // Gauss-Seidel passes the slots of this wave need: the depth of the
  // deepest corner in the upstream order (after the first solve, a pass
  // finalises every corner whose upstream corners are final).
  int passes = 0;
  {
    unsigned long long done = __ballot(need == 0);
    while (done != ~0ull && passes < 8) {
      const unsigned fin = (unsigned)(done >> slotBase) & 0xffu;
      done = __ballot((need & ~fin) == 0);
      ++passes;
    }
  }
  // A slot's record is written and read by the lanes of one wave, whose LDS
  // operations complete in the order they are issued: a wave-level barrier
  // (no code motion across it) orders the reads after the writes.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // Upwind psi of the incoming FP faces (SweepUCBxyz.F90:139-161); a face
  // that brings none reads a zero.
  Real psf[8][3];
  static_for<8>([&](auto C) {
    constexpr int c = C;
    static_for<3>([&](auto F) {
      constexpr int f = F;
      psf[c][f] = psiA[R.off[c][f] + lane];
    });
  });
  // The reciprocals need no upwind psi: in FP64 (four waves per SIMD either
  // way) they are formed while its loads are in flight; FP32 forms them
  // last, which keeps it at 8 waves.
  Real rden[8];
  if constexpr (sizeof(Real) == 8)
    static_for<8>([&](auto C) {
      constexpr int c = C;
      rden[c] = pinned(quot<Real>(1, R.outArea[c] + sig * R.vol[c]));
    });
  // The slot's record is re-read from LDS in each phase below rather than
  // held in registers across phases.
  reread_lds();

  // Vertical edges (i, i + 4): the term at the upwind corner (sez_term with
  // that corner's operands). Computed first: it is the last use of face 0's
  // upwind psi but one, and 4 terms take fewer registers than 8 values.
  Real vt[4];
  static_for<4>([&](auto I) {
    constexpr int i = I, j = I + 4;
    const bool up = (pinned(R.flags[i]) & 2) != 0;
    const Real QU = up ? Q[i] : Q[j], QD = up ? Q[j] : Q[i];
    const Real pU = up ? psf[i][0] : psf[j][0];
    const Real vU = R.vertV[i];
    const Real tU = sig * vU;
    vt[i] = pinned(sez_term<Real>(R.vertA[i], R.vertM[i], tU, tU * tU, sig, vU, QU, QD, pU));
    one_at_a_time();
  });
  reread_lds();
  // Volume and incoming FP terms (FP64 takes the first product off V q in
  // one fused step).
  Real src[8];
  static_for<8>([&](auto C) {
    constexpr int c = C;
    const Real v = R.vol[c], q = Q[c];
    Real s;
    if constexpr (sizeof(Real) == 8) {
#pragma clang fp contract(off)
      s = fma(v, q, -(R.afpIn[c][0] * psf[c][0]));
      s = s - R.afpIn[c][1] * psf[c][1];
      s = s - R.afpIn[c][2] * psf[c][2];
    } else {
      s = v * q;
      static_for<3>([&](auto F) {
        constexpr int f = F;
        s = s - R.afpIn[c][f] * psf[c][f];
      });
    }
    src[c] = pinned(s);
    one_at_a_time();
  });
  reread_lds();
  // Ring edges: corner c owns the edge to its slot-1 neighbour n
  // (ring_term with the upwind corner's operands); the owner adds the term
  // when upwind and takes it off when not, and so does n with its own sign.
  // In each ring the owners go round in order, so a term is added to its
  // owner at once and to n (whose own term comes first) one edge later:
  // at most two terms are held.
  auto ring = [&](auto C) {
    constexpr int c = C, n = kEZ(C, 1);
    const bool up1 = (pinned(R.flags[c]) & 1) != 0;
    const Real w1 = sig * psf[c][2] - Q[c];
    const Real wN = sig * psf[n][1] - Q[n];
    const Real vU = R.ringV[c];
    const Real tU = sig * vU;
    const Real e = pinned(ring_term<Real>(R.ringA[c], R.ringM[c], tU, tU * tU, sig, vU,
                                          up1 ? w1 : wN, Q[c] - Q[n]));
    src[c] = fma(R.sg1[c], e, src[c]);
    one_at_a_time();
    return e;
  };
  static_for<2>([&](auto H) {
    // Ring H: corners 4H .. 4H + 3 in the order each owns the edge its
    // slot-0 neighbour takes from it (bottom 0 1 2 3, top 4 7 6 5).
    constexpr int r0 = 4 * H, r1 = H ? 7 : 1, r2 = H ? 6 : 2, r3 = H ? 5 : 3;
    const Real e0 = ring(std::integral_constant<int, r0>{});
    const Real e1 = ring(std::integral_constant<int, r1>{});
    src[r1] = fma(R.sg0[r1], e0, src[r1]);
    const Real e2 = ring(std::integral_constant<int, r2>{});
    src[r2] = fma(R.sg0[r2], e1, src[r2]);
    const Real e3 = ring(std::integral_constant<int, r3>{});
    src[r3] = fma(R.sg0[r3], e2, src[r3]);
    src[r0] = fma(R.sg0[r0], e3, src[r0]);
  });
  reread_lds();
  static_for<8>([&](auto C) {
    constexpr int c = C;
    src[c] = fma(R.vertS[c], vt[c & 3], src[c]);
  });
  reread_lds();

  // Corner solve psi_c = (src_c + sum over upstream EZ neighbours of
  // |aez| psi_n) / den_c, den_c = outgoing areas + sigma V_c.
  Real psi[8];
  static_for<8>([&](auto C) {
    constexpr int c = C;
    if constexpr (sizeof(Real) == 4) rden[c] = quot<Real>(1, R.outArea[c] + sig * R.vol[c]);
    psi[c] = src[c] * rden[c];
  });
  if (cyclic) {
    // A cyclic zone is solved once from the previous values of its own
    // corners (SweepUCBxyz.F90:281-295), which psi holds: its corners are on
    // the angle's cycle list. Its passes then leave it as it is (no
    // upstream terms, unit reciprocal).
    Real old[8];
    static_for<8>([&](auto C) {
      constexpr int c = C;
      old[c] = psiA[(8L * z + c) * G + g];
    });
    static_for<8>([&](auto C) {
      constexpr int c = C;
      psi[c] = (src[c] + R.kOnce[c][0] * old[kEZ(C, 0)] + R.kOnce[c][1] * old[kEZ(C, 1)] +
                R.kOnce[c][2] * old[kEZ(C, 2)]) * rden[c];
    });
    static_for<8>([&](auto C) {
      constexpr int c = C;
      src[c] = psi[c];
      rden[c] = 1;
    });
  }
  for (int pass = 0; pass < passes; ++pass) {
    reread_lds();
    static_for<8>([&](auto C) {
      constexpr int c = C;
      psi[c] = (src[c] + R.kup[c][0] * psi[kEZ(C, 0)] + R.kup[c][1] * psi[kEZ(C, 1)] +
                R.kup[c][2] * psi[kEZ(C, 2)]) * rden[c];
    });
  }

  reread_lds();
  if (valid) {
    static_for<8>([&](auto C) {
      constexpr int c = C;
      psiA[(8L * z + c) * G + g] = psi[c];
    });
    static_for<8>([&](auto C) {
      constexpr int c = C;
      const int fl = R.flags[c];
      if (fl & 28)
        static_for<3>([&](auto F) {
          constexpr int f = F;
          if (fl & (4 << f)) psiA[(long)R.cfp[c][f] * G + g] = psi[c];
        });
    });
  }
}

int zone_lanes_for(int G) {
  return G % 64 == 0 ? 64 : (G % 32 == 0 ? 32 : (G % 16 == 0 ? 16 : 0));
}

// One level of many domains in one launch. The level's items are grouped by
// domain into segments; segment s starts at task segs[s].taskStart, where a
// task is one (item, K groups) pair of 8 lanes. Segments are padded to whole
// waves, so a wave finds its segment once, by a binary search on its first
// task with wave-uniform (scalar) loads, takes its domain's arrays the same
// way, and runs the same corner code as the one-domain kernel; padding tasks
// compute the segment's first item and store nothing.
template <typename Real, int K>
__device__ __forceinline__ void batch_task(const BatchLevelArgs<Real>& b, unsigned task) {
  const unsigned waveTask = __builtin_amdgcn_readfirstlane(task);
  int lo = 0, hi = b.nSegs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b.segs[mid].taskStart <= (long)waveTask) lo = mid;
    else hi = mid - 1;
  }
  const BatchSeg sg = b.segs[lo];
  const BatchDomain<Real>& d = b.doms[sg.domain];
  const unsigned inSeg = (unsigned)((long)task - sg.taskStart);
  const bool valid = (long)task < b.nTasks && inSeg < (unsigned)sg.nTasks;
  SweepLevelArgs<Real> p{};
  p.items = b.items;
  p.G = d.G;
  p.C = d.C;
  p.Afp = d.Afp;
  p.Aez = d.Aez;
  p.vol = d.vol;
  p.cFP = d.cFP;
  p.S = d.S;
  p.sigt = d.sigt;
  p.omega = d.omega;
  p.psi = d.psi;
  p.psiAngleStride = d.psiAngleStride;
  p.angle0 = b.angle0;
  p.parTol = 0;
  const unsigned blocksPerItem = (unsigned)(d.G / K);
  const unsigned local = valid ? inSeg : 0u;
  const unsigned q = fast_div(local, sg.perItem);
  const int item = sg.itemStart + (int)q;
  const int g0 = K * (int)(local - q * blocksPerItem);
  sweep_corner<Real, K, false>(p, item, g0, valid);
}

template <typename Real, int K>
__global__ __launch_bounds__(256) void ucb_sweep_batch_level(BatchLevelArgs<Real> b) {
  batch_task<Real, K>(b, (blockIdx.x * blockDim.x + threadIdx.x) >> 3);
}

// Grid-wide barrier of the persistent batch kernel. Every workgroup
// publishes its writes (agent-scope release) and arrives on the counter of
// its group (workgroups w with the same w mod 8); the last of a group
// arrives on the top counter, and the last there resets it and starts the
// next generation, which the others wait for (acquire). Arrivals on eight
// lines, not one, keep them from queueing on a single address. The wait is
// bounded: after spinLimit polls (0: at once), or when another workgroup
// has already given up, the workgroup sets the error flag and returns
// false, so every wave leaves the kernel whether or not the whole grid was
// resident. The host checks the flag.
__device__ __forceinline__ bool grid_barrier(GridBarrier* bar, unsigned spinLimit) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    __threadfence();
    const unsigned gen = __hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned grp = blockIdx.x & 7u;
    const unsigned inGroup = (gridDim.x - grp + 7u) >> 3;   // workgroups w with w mod 8 = grp
    const unsigned groups = gridDim.x < 8u ? gridDim.x : 8u;
    bool last = false;
    if (__hip_atomic_fetch_add(&bar->sub[grp][0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u ==
        inGroup) {
      __hip_atomic_store(&bar->sub[grp][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u ==
             groups;
    }
    if (last) {
      __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&bar->gen, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned polls = 0;
      while (__hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (polls++ >= spinLimit ||
            __hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
          __hip_atomic_store(&bar->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    }
    __threadfence();
  }
  __syncthreads();
  return ok != 0;
}

// Every level of one angle block of a batch in one launch: a resident grid
// walks the levels in order, each workgroup taking kPersistentThreads / 8
// tasks at a time with a grid stride, with a grid barrier between levels.
// The per-task work is the per-level kernel's, so both give the same bits.
template <typename Real, int K>
__global__ __launch_bounds__(kPersistentThreads) void ucb_sweep_batch_persistent(
    BatchPersistentArgs<Real> b) {
  constexpr long kTasks = kPersistentThreads / 8;
  for (int h = b.h0; h < b.h1; ++h) {
    BatchLevelArgs<Real> a{};
    a.doms = b.doms;
    a.items = b.items + b.levelItem[h];
    a.segs = b.segs + b.levelSeg[h];
    a.nSegs = b.levelSeg[h + 1] - b.levelSeg[h];
    a.nTasks = b.levelTasks[h];
    a.angle0 = b.angle0;
    for (long base = kTasks * blockIdx.x; base < a.nTasks; base += kTasks * gridDim.x)
      batch_task<Real, K>(a, (unsigned)(base + (threadIdx.x >> 3)));
    if (h + 1 < b.h1 && !grid_barrier(b.bar, b.spinLimit)) break;
  }
}

// Incident boundary fluxes into the boundary tails of an angle block's psi
// (the caller's PsiB(:, :, Angle) loaded into Psi1's tail,
// snac/SweepUCBxyz.F90:95-97). One thread per (angle, boundary element, group).
template <typename Real>
__global__ __launch_bounds__(256) void boundary_in(const Real* __restrict__ inc, long nbG, long CG,
                                                   int a0, int nA, long angleStride,
                                                   Real* __restrict__ psi) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbG * nA) return;
  const long a = i / nbG;
  const long k = i - a * nbG;
  psi[a * angleStride + CG + k] = inc[(a0 + a) * nbG + k];
}

// Cycle-list corners (control/constructDynMemory.F90:118-189): entry m is
// (corner, angle), 0-based; toPsi copies cyclePsi(:, m) into psi before the
// angle block is swept (initFromCycleList), otherwise psi into cyclePsi after
// it (updateCycleList). Entries first .. first + n - 1, one thread per group.
template <typename Real>
__global__ __launch_bounds__(256) void cycle_copy(const int2* __restrict__ entries, long first,
                                                  long n, int G, Real* __restrict__ cyclePsi,
                                                  Real* __restrict__ psi, long angleStride,
                                                  int angle0, int toPsi) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * G) return;
  const long m = first + i / G;
  const int g = (int)(i - (i / G) * G);
  const int2 e = entries[m];
  Real* p = psi + (long)(e.y - angle0) * angleStride + (long)e.x * G + g;
  Real* q = cyclePsi + m * G + g;
  if (toPsi) *p = *q;
  else *q = *p;
}

// phi(g,c) = sum_a w_a psi(g,c,a), accumulated in FP64 in angle order
// (or added to phi when accumulate is set).
template <typename Real>
__global__ __launch_bounds__(256) void phi_from_psi(const Real* __restrict__ psi,
                                                    const double* __restrict__ weight, int A,
                                                    long angleStride, long n,
                                                    double* __restrict__ phi, int accumulate) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // Continuing from phi when accumulating keeps the sum in angle order, so
  // sweeping the angles in blocks gives the same bits as one block.
  double acc = accumulate ? phi[i] : 0.0;
  for (int a = 0; a < A; ++a) acc += weight[a] * (double)psi[(long)a * angleStride + i];
  phi[i] = acc;
}

// The same sum for V adjacent elements per thread with 16-byte loads (V = 2
// for FP64 psi, 4 for FP32), angles unrolled so several loads are in flight;
// each element's sum runs in angle order as above, so the bits are the same.
// Needs angleStride and n multiples of V and 16-byte aligned psi and phi.
template <typename Real, int V>
__global__ __launch_bounds__(256) void phi_from_psi_vec(const Real* __restrict__ psi,
                                                        const double* __restrict__ weight, int A,
                                                        long angleStride, long n,
                                                        double* __restrict__ phi, int accumulate) {
  using Vec = GroupVec<Real, V>;
  const long i = V * ((long)blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  double acc[V];
#pragma unroll
  for (int k = 0; k < V; ++k) acc[k] = accumulate ? phi[i + k] : 0.0;
  // psi is read once here: non-temporal loads keep it from displacing
  // reusable lines in L2 and the MALL.
  typedef Real Raw __attribute__((ext_vector_type(V)));
#pragma unroll 8
  for (int a = 0; a < A; ++a) {
    const Raw x = __builtin_nontemporal_load((const Raw*)(psi + (long)a * angleStride + i));
    const double w = weight[a];
#pragma unroll
    for (int k = 0; k < V; ++k) acc[k] += w * (double)x[k];
  }
#pragma unroll
  for (int k = 0; k < V; ++k) phi[i + k] = acc[k];
}

// Source for the next sweep: S = (sigma_s phi + q) / (4 pi), within-group
// isotropic scattering (getCollisionRate semantics for a steady problem).
template <typename Real>
__global__ __launch_bounds__(256) void source_update(const double* __restrict__ phi,
                                                     const double* __restrict__ sigs,
                                                     const double* __restrict__ q, int G,
                                                     long n, Real* __restrict__ S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c = i / G;
  const int g = (int)(i - c * G);
  const long zg = (c >> 3) * G + g;
  const double wtiso = 0.07957747154594767;  // 1/(4 pi)
  S[i] = (Real)(wtiso * (sigs[zg] * phi[i] + q[zg]));
}

// Zone averages phibar(g,z) = sum_c V_c phi(g,c) / V_z and the running max
// (of the values, or of their magnitudes when magnitude != 0).
__global__ __launch_bounds__(256) void zone_average(const double* __restrict__ phi,
                                                    const double* __restrict__ vol,
                                                    const double* __restrict__ volZ, int G,
                                                    long n, double* __restrict__ phibar,
                                                    unsigned long long* __restrict__ maxBits,
                                                    int magnitude) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const long z = i / G;
    const int g = (int)(i - z * G);
    double acc = 0.0;
    for (int c = 0; c < 8; ++c) acc += phi[(8 * z + c) * G + g] * vol[8 * z + c];
    v = acc / volZ[z];
    phibar[i] = v;
  }
  // Wave-level max of non-negative values, then one atomic per wave. A
  // non-finite average (NaN, or +-Inf) becomes NaN and is kept, so the
  // residual that follows is NaN and the host stops the iteration.
  if (magnitude) v = fabs(v);
  double m = v > 0.0 ? v : 0.0;
  if (!isfinite(v)) m = __longlong_as_double((long long)kNaNBits);
  for (int off = 32; off > 0; off >>= 1) m = nan_keep_max(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0 && (m > 0.0 || m != m))
    atomicMax(maxBits, (unsigned long long)__double_as_longlong(m));
}

// Max relative change of zone averages over entries above cutoff * max
// (rt/ConvergenceTest.F90:150-168 semantics per zone and group), and copy of
// the new averages into the old. With magnitude != 0 the test is on |phibar|
// (the same for non-negative fluxes; corrected iterates may go negative).
__global__ __launch_bounds__(256) void residual_max(const double* __restrict__ phibar,
                                                    double* __restrict__ phibarOld, long n,
                                                    const unsigned long long* __restrict__ maxBits,
                                                    double cutoff,
                                                    unsigned long long* __restrict__ resBits,
                                                    int magnitude) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (i < n) {
    const double mx = __longlong_as_double((long long)*maxBits);
    const double thr = cutoff * mx;
    const double v = phibar[i];
    const double a = magnitude ? fabs(v) : v;
    if (a > thr) r = fabs((v - phibarOld[i]) / v);
    // A non-finite average anywhere (here, or behind a non-finite max)
    // makes the residual NaN.
    if (!isfinite(v) || !isfinite(mx)) r = __longlong_as_double((long long)kNaNBits);
    phibarOld[i] = v;
  }
  for (int off = 32; off > 0; off >>= 1) r = nan_keep_max(r, __shfl_down(r, off, 64));
  if ((threadIdx.x & 63) == 0 && (r > 0.0 || r != r))
    atomicMax(resBits, (unsigned long long)__double_as_longlong(r));
}

__global__ void fill_u64(unsigned long long* p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void fill_zero(T* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = T(0);
}

// Streaming copy used only to calibrate HBM byte counters.
__global__ __launch_bounds__(256) void umt_copy_probe(const double* __restrict__ s,
                                                      double* __restrict__ d, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) d[i] = s[i];
}

template <typename Dst, typename Src>
__global__ __launch_bounds__(256) void convert_copy(const Src* __restrict__ s, Dst* __restrict__ d,
                                                    long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (Dst)s[i];
}

// ----------------------------------------------------------------------------
// Launch wrappers (host side).

static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename Real, int K>
static hipError_t launch_k(const SweepLevelArgs<Real>& a, bool singleAngle, hipStream_t s) {
  const long threads = 8L * a.nItems * (a.G / K);  // one lane per corner of each (item, K groups)
  if (threads == 0) return hipSuccess;
  if (threads > kMaxLevelThreads) return hipErrorInvalidValue;
  SweepLevelArgs<Real> b = a;
  b.perItem = make_fast_div((unsigned)(a.G / K));
  if (singleAngle)
    hipLaunchKernelGGL((ucb_sweep_level<Real, K, true>), dim3(blocks_for(threads, 256)), dim3(256),
                       0, s, b);
  else
    hipLaunchKernelGGL((ucb_sweep_level<Real, K, false>), dim3(blocks_for(threads, 256)),
                       dim3(256), 0, s, b);
  return hipGetLastError();
}

int groups_per_lane_supported(int G, int K) {
  return (K == 1 || K == 2 || K == 4 || K == 8) && G % K == 0;
}

template <typename Real>
hipError_t launch_sweep_batch_persistent(const BatchPersistentArgs<Real>& b, int K, int blocks,
                                         hipStream_t s) {
  if (blocks < 1) return hipErrorInvalidValue;
  const dim3 grid(blocks), block(kPersistentThreads);
  switch (K) {
    case 4: hipLaunchKernelGGL((ucb_sweep_batch_persistent<Real, 4>), grid, block, 0, s, b); break;
    case 2: hipLaunchKernelGGL((ucb_sweep_batch_persistent<Real, 2>), grid, block, 0, s, b); break;
    default: hipLaunchKernelGGL((ucb_sweep_batch_persistent<Real, 1>), grid, block, 0, s, b); break;
  }
  return hipGetLastError();
}

template <typename Real, int K>
static int resident_blocks_k() {
  int perCU = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, ucb_sweep_batch_persistent<Real, K>,
                                                    kPersistentThreads, 0) !=
          hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return perCU * cus;
}

int batch_persistent_resident_blocks(int precision, int K) {
  if (precision == 32)
    return K == 4 ? resident_blocks_k<float, 4>() : K == 2 ? resident_blocks_k<float, 2>()
                                                           : resident_blocks_k<float, 1>();
  return K == 4 ? resident_blocks_k<double, 4>() : K == 2 ? resident_blocks_k<double, 2>()
                                                          : resident_blocks_k<double, 1>();
}

template <typename Real>
hipError_t launch_sweep_batch_level(const BatchLevelArgs<Real>& b, int K, hipStream_t s) {
  const long threads = 8L * b.nTasks;
  if (threads == 0) return hipSuccess;
  if (threads > kMaxLevelThreads) return hipErrorInvalidValue;
  const dim3 grid(blocks_for(threads, 256)), block(256);
  switch (K) {
    case 4: hipLaunchKernelGGL((ucb_sweep_batch_level<Real, 4>), grid, block, 0, s, b); break;
    case 2: hipLaunchKernelGGL((ucb_sweep_batch_level<Real, 2>), grid, block, 0, s, b); break;
    default: hipLaunchKernelGGL((ucb_sweep_batch_level<Real, 1>), grid, block, 0, s, b); break;
  }
  return hipGetLastError();
}

template <typename Real, int W>
static hipError_t launch_zone(const SweepLevelArgs<Real>& a, hipStream_t s) {
  const long threads = (long)a.nItems * a.G;  // one lane per (item, group)
  if (threads == 0) return hipSuccess;
  if (threads > kMaxLevelThreads) return hipErrorInvalidValue;
  SweepLevelArgs<Real> b = a;
  b.perItem = make_fast_div((unsigned)(a.G / W));
  hipLaunchKernelGGL((ucb_sweep_level_zone<Real, W>), dim3(blocks_for(threads, 256)), dim3(256), 0,
                     s, b);
  return hipGetLastError();
}

int sweep_layout_supported(int G, int K) {
  return K == kZoneLayout ? zone_lanes_for(G) != 0 : groups_per_lane_supported(G, K);
}

template <typename Real>
hipError_t launch_sweep_level(const SweepLevelArgs<Real>& a, bool singleAngle, hipStream_t s) {
  if (a.groupsPerLane == kZoneLayout) {
    // The zone layout serves all-angle sweeps with G a multiple of 16; the
    // per-angle drop-in and other G run the corner layout's widest K.
    if (!singleAngle) switch (zone_lanes_for(a.G)) {
        case 64: return launch_zone<Real, 64>(a, s);
        case 32: return launch_zone<Real, 32>(a, s);
        case 16: return launch_zone<Real, 16>(a, s);
        default: break;
      }
    const int K = a.G % 8 == 0 ? 8 : (a.G % 4 == 0 ? 4 : (a.G % 2 == 0 ? 2 : 1));
    SweepLevelArgs<Real> b = a;
    b.groupsPerLane = K;
    return launch_sweep_level<Real>(b, singleAngle, s);
  }
  switch (a.groupsPerLane) {
    case 8: return launch_k<Real, 8>(a, singleAngle, s);
    case 4: return launch_k<Real, 4>(a, singleAngle, s);
    case 2: return launch_k<Real, 2>(a, singleAngle, s);
    default: return launch_k<Real, 1>(a, singleAngle, s);
  }
}

template <typename Real>
hipError_t launch_phi_from_psi_acc(const Real* psi, const double* w, int A, long stride, long n,
                                   double* phi, bool accumulate, hipStream_t s) {
  constexpr int V = 16 / sizeof(Real);
  if (stride % V == 0 && n % V == 0 && (uintptr_t)psi % 16 == 0 && (uintptr_t)phi % 16 == 0)
    hipLaunchKernelGGL((phi_from_psi_vec<Real, V>), dim3(blocks_for(n / V, 256)), dim3(256), 0, s,
                       psi, w, A, stride, n, phi, accumulate ? 1 : 0);
  else
    hipLaunchKernelGGL((phi_from_psi<Real>), dim3(blocks_for(n, 256)), dim3(256), 0, s, psi, w, A,
                       stride, n, phi, accumulate ? 1 : 0);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_phi_from_psi(const Real* psi, const double* w, int A, long stride, long n,
                               double* phi, hipStream_t s) {
  return launch_phi_from_psi_acc<Real>(psi, w, A, stride, n, phi, false, s);
}

template <typename Real>
hipError_t launch_source_update(const double* phi, const double* sigs, const double* q, int G,
                                long n, Real* S, hipStream_t s) {
  hipLaunchKernelGGL((source_update<Real>), dim3(blocks_for(n, 256)), dim3(256), 0, s, phi, sigs,
                     q, G, n, S);
  return hipGetLastError();
}

hipError_t launch_residual(const double* phi, const double* vol, const double* volZ, int G, long nzg,
                           double* phibar, double* phibarOld, double cutoff,
                           unsigned long long* bits, hipStream_t s, bool magnitude) {
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, s, bits, 2, 0ull);
  hipLaunchKernelGGL(zone_average, dim3(blocks_for(nzg, 256)), dim3(256), 0, s, phi, vol, volZ, G,
                     nzg, phibar, bits, (int)magnitude);
  hipLaunchKernelGGL(residual_max, dim3(blocks_for(nzg, 256)), dim3(256), 0, s, phibar, phibarOld,
                     nzg, bits, cutoff, bits + 1, (int)magnitude);
  return hipGetLastError();
}

hipError_t launch_zone_average(const double* phi, const double* vol, const double* volZ, int G,
                               long nzg, double* phibar, unsigned long long* maxBits,
                               hipStream_t s) {
  hipLaunchKernelGGL(zone_average, dim3(blocks_for(nzg, 256)), dim3(256), 0, s, phi, vol, volZ, G,
                     nzg, phibar, maxBits, 0);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_zero(T* p, long n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((fill_zero<T>), dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_copy_probe(const double* src, double* dst, long n, hipStream_t s) {
  const long blocks = std::min<long>(blocks_for(n, 256), 256L * 32);
  hipLaunchKernelGGL(umt_copy_probe, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_cycle_copy(const int2* entries, long first, long n, int G, Real* cyclePsi,
                             Real* psi, long angleStride, int angle0, bool toPsi, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((cycle_copy<Real>), dim3(blocks_for(n * G, 256)), dim3(256), 0, s, entries,
                     first, n, G, cyclePsi, psi, angleStride, angle0, toPsi ? 1 : 0);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_boundary_in(const Real* inc, long nbG, long CG, int a0, int nA, long angleStride,
                              Real* psi, hipStream_t s) {
  if (nbG <= 0 || nA <= 0) return hipSuccess;
  hipLaunchKernelGGL((boundary_in<Real>), dim3(blocks_for(nbG * nA, 256)), dim3(256), 0, s, inc, nbG,
                     CG, a0, nA, angleStride, psi);
  return hipGetLastError();
}

template <typename Dst, typename Src>
hipError_t launch_convert(const Src* s_, Dst* d, long n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((convert_copy<Dst, Src>), dim3(blocks_for(n, 256)), dim3(256), 0, s, s_, d, n);
  return hipGetLastError();
}

template hipError_t launch_sweep_batch_level<double>(const BatchLevelArgs<double>&, int, hipStream_t);
template hipError_t launch_sweep_batch_level<float>(const BatchLevelArgs<float>&, int, hipStream_t);
template hipError_t launch_sweep_batch_persistent<double>(const BatchPersistentArgs<double>&, int,
                                                         int, hipStream_t);
template hipError_t launch_sweep_batch_persistent<float>(const BatchPersistentArgs<float>&, int, int,
                                                        hipStream_t);
template hipError_t launch_sweep_level<double>(const SweepLevelArgs<double>&, bool, hipStream_t);
template hipError_t launch_sweep_level<float>(const SweepLevelArgs<float>&, bool, hipStream_t);
template hipError_t launch_phi_from_psi<double>(const double*, const double*, int, long, long,
                                                double*, hipStream_t);
template hipError_t launch_phi_from_psi<float>(const float*, const double*, int, long, long, double*,
                                               hipStream_t);
template hipError_t launch_phi_from_psi_acc<double>(const double*, const double*, int, long, long,
                                                    double*, bool, hipStream_t);
template hipError_t launch_phi_from_psi_acc<float>(const float*, const double*, int, long, long,
                                                   double*, bool, hipStream_t);
template hipError_t launch_source_update<double>(const double*, const double*, const double*, int,
                                                 long, double*, hipStream_t);
template hipError_t launch_source_update<float>(const double*, const double*, const double*, int,
                                                long, float*, hipStream_t);
template hipError_t launch_fill_zero<double>(double*, long, hipStream_t);
template hipError_t launch_fill_zero<float>(float*, long, hipStream_t);
template hipError_t launch_convert<float, double>(const double*, float*, long, hipStream_t);
template hipError_t launch_convert<double, float>(const float*, double*, long, hipStream_t);
template hipError_t launch_convert<float, float>(const float*, float*, long, hipStream_t);
template hipError_t launch_boundary_in<double>(const double*, long, long, int, int, long, double*,
                                               hipStream_t);
template hipError_t launch_boundary_in<float>(const float*, long, long, int, int, long, float*,
                                              hipStream_t);
template hipError_t launch_cycle_copy<double>(const int2*, long, long, int, double*, double*, long,
                                              int, bool, hipStream_t);
template hipError_t launch_cycle_copy<float>(const int2*, long, long, int, float*, float*, long,
                                             int, bool, hipStream_t);
template hipError_t launch_convert<double, double>(const double*, double*, long, hipStream_t);

}  // namespace umt
