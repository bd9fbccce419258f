// CDNA4 kernels for the 2D cylindrical (RZ) upstream-corner-balance sweep.
//
// One lane owns one corner of one (zone, direction) item for K consecutive
// groups: the four corners of a quad zone are four consecutive lanes, so
// their EZ neighbours (the next and previous corner of the clockwise quad)
// are one quad_perm DPP move away. A level holds (zone, direction) items of
// every xi level whose upwind zones and whose previous direction at the
// same zone are done, so the directions of one xi level pipeline through
// the mesh one level apart and all xi levels run side by side.
//
// The corner arithmetic follows snac/SweepUCBrz.F90:130-276: radius-weighted
// incoming FP fluxes, the EZ coupling sez with fouralpha = 1.82 on the
// corner area, the angular-derivative term fac * Area * PsiM with the
// half-angle flux PsiM carried from the previous direction, and the corner
// solve psi_c = (src_c + fac A_c PsiM_c) / (sumArea_c + sigma V_c), where
// sumArea_c = fac A_c + the radius-weighted incoming areas. The within-zone
// solve relaxes all four corners at once until each has seen final upstream
// values, as the xyz kernel does.
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "kernels.h"
#include "lane_moves.h"

namespace umt {

namespace {

// Value of this lane's EZ neighbour across corner face f: face 1 (f = 0)
// leads to the next corner of the quad, face 2 (f = 1) to the previous one.
template <int F, typename T>
__device__ __forceinline__ T rz_from_ez(T x) {
  return F == 0 ? dpp_move<kQuadNext>(x) : dpp_move<kQuadPrev>(x);
}

template <typename Real>
__device__ __forceinline__ Real rz_sez(Real R, Real A, Real sig, Real area, Real Qu, Real Qd,
                                       bool withOpp, Real psiOpp) {
  if (withOpp) {
    const Real sigA = sig * area;
    const Real sigA2 = sigA * sigA;
    const Real gnum = A * A * ((Real)1.82 * sigA2 + A * ((Real)4 * sigA + (Real)3 * A));
    const Real gden =
        area * ((Real)4 * sigA * sigA2 + A * ((Real)6 * sigA2 + (Real)2 * A * ((Real)2 * sigA + A)));
    return quot<Real>(R * (area * gnum * (sig * psiOpp - Qu) + (Real)0.5 * A * gden * (Qu - Qd)),
                      gnum + gden * sig);
  }
  return quot<Real>((Real)0.5 * R * A * (Qu - Qd), sig);
}

}  // namespace

// One lane: corner c of zone z in direction a for groups g0 .. g0 + K - 1;
// the zone's four corners are four consecutive lanes of a quad. Lanes with
// valid == false compute but store nothing.
template <typename Real, int K>
__device__ __forceinline__ void rz_corner(const RzLevelArgs<Real>& p, int z, int a, int c, int g0,
                                          bool valid) {
  const int G = p.G;
  const RzAngle an = p.angles[a];
  const bool starting = an.flags & 1, finishingNext = an.flags & 2;
  const long cc = 4L * z + c;
  const Real mu = (Real)an.mu, xi = (Real)an.xi;
  Real* __restrict__ psiA = p.psi + (long)a * p.psiAngleStride;

  Real afp[2], aez[2], rfp[2], rez[2];
  int cfp[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    afp[f] = mu * p.Afp[4 * cc + 2 * f] + xi * p.Afp[4 * cc + 2 * f + 1];
    aez[f] = mu * p.Aez[4 * cc + 2 * f] + xi * p.Aez[4 * cc + 2 * f + 1];
    rfp[f] = p.Rfp[2 * cc + f];
    rez[f] = p.Rez[2 * cc + f];
    cfp[f] = p.cFP[2 * cc + f] - 1;
  }
  const Real area = p.area[cc], v = p.vol[cc], fac = (Real)an.fac;
  // sumArea: fac A + radius-weighted incoming FP and EZ areas (SweepUCBrz.F90:128,147,161).
  Real inArea = fac * area;
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    if (afp[f] < 0) inArea -= rfp[f] * afp[f];
    if (aez[f] < 0) inArea -= rez[f] * aez[f];
  }
  const GroupVec<Real, K> sg = load_groups<Real, K>(p.sigt + (long)z * G + g0);
  GroupVec<Real, K> Q = load_groups<Real, K>(p.S + cc * G + g0);
  if (p.tauPsi) {  // time-dependent source STotal + tau Psi (SweepUCBrz.F90:136)
    const GroupVec<Real, K> prev =
        load_groups<Real, K>(p.tauPsi + ((long)a * p.C + cc) * G + g0);
#pragma unroll
    for (int k = 0; k < K; ++k) Q.v[k] += p.tau * prev.v[k];
  }
  // Incoming FP fluxes; boundaries are vacuum in these sweeps.
  GroupVec<Real, K> psf[2];
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    if (afp[f] < 0 && cfp[f] < p.C) {
      psf[f] = load_groups<Real, K>(psiA + (long)cfp[f] * G + g0);
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) psf[f].v[k] = 0;
    }
  }
  // Half-angle flux from the previous direction of the xi level (none for
  // a starting direction, whose fac is zero).
  Real* __restrict__ pmA = p.psiM + (long)an.level * p.C * G;
  GroupVec<Real, K> pm;
  if (starting) {
#pragma unroll
    for (int k = 0; k < K; ++k) pm.v[k] = 0;
  } else {
    pm = load_groups<Real, K>(pmA + cc * G + g0);
  }
  Real src[K], rden[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const Real sig = sg.v[k], q = Q.v[k];
    Real s = v * q;
#pragma unroll
    for (int f = 0; f < 2; ++f) s -= afp[f] < 0 ? rfp[f] * afp[f] * psf[f].v[k] : (Real)0;
    const Real Qn0 = rz_from_ez<0>(q), Qn1 = rz_from_ez<1>(q);
    const Real sez0 = aez[0] > 0 ? rz_sez<Real>(rez[0], aez[0], sig, area, q, Qn0, afp[0] < 0, psf[0].v[k])
                                 : (Real)0;
    const Real sez1 = aez[1] > 0 ? rz_sez<Real>(rez[1], aez[1], sig, area, q, Qn1, afp[1] < 0, psf[1].v[k])
                                 : (Real)0;
    s += sez0;
    s += sez1;
    // The next corner sees this shared face as its face 2, the previous one as its face 1.
    s -= rz_from_ez<0>(sez1);
    s -= rz_from_ez<1>(sez0);
    s += area * fac * pm.v[k];
    src[k] = s;
    rden[k] = quot<Real>(1, inArea + sig * v);
  }
  // Corner solve: psi_c = (src_c + sum over upwind EZ neighbours of R |aez| psi_n) / den.
  const Real k0 = aez[0] < 0 ? -rez[0] * aez[0] : (Real)0;
  const Real k1 = aez[1] < 0 ? -rez[1] * aez[1] : (Real)0;
  GroupVec<Real, K> psi;
#pragma unroll
  for (int k = 0; k < K; ++k) psi.v[k] = src[k] * rden[k];
  int fin = (k0 == 0 && k1 == 0) ? 1 : 0;
  for (int iter = 0; iter < 4; ++iter) {
    if (__all(fin)) break;
    const int f0 = rz_from_ez<0>(fin), f1 = rz_from_ez<1>(fin);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Real x = psi.v[k];
      psi.v[k] = (src[k] + k0 * rz_from_ez<0>(x) + k1 * rz_from_ez<1>(x)) * rden[k];
    }
    fin = (k0 == 0 || f0) && (k1 == 0 || f1);
  }
  // Next half-angle flux (SweepUCBrz.F90:233-245).
  GroupVec<Real, K> pmNew;
#pragma unroll
  for (int k = 0; k < K; ++k)
    pmNew.v[k] = starting ? psi.v[k] : (Real)an.w1 * psi.v[k] - (Real)an.w2 * pm.v[k];
  if (valid) {
    store_groups<Real, K>(psiA + cc * G + g0, psi);
    store_groups<Real, K>(pmA + cc * G + g0, pmNew);
#pragma unroll
    for (int f = 0; f < 2; ++f)
      if (afp[f] > 0 && cfp[f] >= p.C) store_groups<Real, K>(psiA + (long)cfp[f] * G + g0, psi);
    if (finishingNext) {
      // The finishing direction is not swept: its psi (and exiting boundary
      // psi) is the last half-angle flux (SweepUCBrz.F90:259-276).
      Real* psiF = psiA + p.psiAngleStride;
      store_groups<Real, K>(psiF + cc * G + g0, pmNew);
#pragma unroll
      for (int f = 0; f < 2; ++f)
        if (afp[f] > 0 && cfp[f] >= p.C) store_groups<Real, K>(psiF + (long)cfp[f] * G + g0, pmNew);
    }
  }
}

template <typename Real, int K>
__global__ __launch_bounds__(256) void rz_sweep_level(RzLevelArgs<Real> p) {
  const unsigned task = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
  const unsigned perItem = (unsigned)(p.G / K);
  int item = (int)fast_div(task, p.perItem);
  const int g0 = K * (int)(task - (unsigned)item * perItem);
  const bool valid = item < p.nItems;
  if (!valid) item = 0;
  const int2 it = p.items[item];
  rz_corner<Real, K>(p, it.x, it.y, threadIdx.x & 3, g0, valid);
}

// One wave per (tile, direction) item of a tile level and chunk of group
// blocks: the tile's zones are swept in the item's steps, the zones of step
// s reading upwind psi written at earlier steps (or by earlier launches),
// with a barrier between steps. The 16 quads of the wave are 16 / gbPerChunk
// zone positions times gbPerChunk group blocks; a step with more zones than
// positions takes several passes. Every lane runs every step's barrier.
// live and the pass count depend on the quad only, so the lane moves inside
// rz_corner, which stay within a quad, never read a switched-off lane.
template <typename Real, int K>
__global__ __launch_bounds__(64) void rz_sweep_tile(RzTileArgs<Real> q) {
  const int item = blockIdx.x;
  const int quad = threadIdx.x >> 2;
  const int pos = quad / q.gbPerChunk;
  const int positions = 16 / q.gbPerChunk;
  const int gb = blockIdx.y * q.gbPerChunk + (quad - pos * q.gbPerChunk);
  const bool live = gb < q.p.G / K;
  const int2 ti = q.items[item];
  const unsigned char* order = q.order + (long)item * q.maxZones;
  const unsigned char* off = q.stepOff + (long)item * (q.maxZones + 1);
  const int* zones = q.tileZone + (long)ti.x * q.maxZones;
  const int nsteps = q.nsteps[item];
  for (int s = 0; s < nsteps; ++s) {
    const int hi = off[s + 1];
    for (int j = off[s] + pos; j < hi; j += positions)
      if (live) rz_corner<Real, K>(q.p, zones[order[j]], ti.y, threadIdx.x & 3, K * gb, true);
    __syncthreads();
  }
}

// S = wtiso (sigma_s phi + q) with wtiso = 1/(2 pi) for RZ.
template <typename Real>
__global__ __launch_bounds__(256) void rz_source_update(const double* __restrict__ phi,
                                                        const double* __restrict__ sigs,
                                                        const double* __restrict__ q, int G,
                                                        long n, Real* __restrict__ S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c = i / G;
  const int g = (int)(i - c * G);
  const long zg = (c >> 2) * G + g;
  const double wtiso = 0.15915494309189535;  // 1/(2 pi)
  S[i] = (Real)(wtiso * (sigs[zg] * phi[i] + q[zg]));
}

// Zone averages over the four corners, their running max, then the max
// relative change (rt/ConvergenceTest.F90:150-168 semantics).
__global__ __launch_bounds__(256) void rz_zone_average(const double* __restrict__ phi,
                                                       const double* __restrict__ vol,
                                                       const double* __restrict__ volZ, int G,
                                                       long n, double* __restrict__ phibar,
                                                       unsigned long long* __restrict__ maxBits) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const long z = i / G;
    const int g = (int)(i - z * G);
    double acc = 0.0;
    for (int c = 0; c < 4; ++c) acc += phi[(4 * z + c) * G + g] * vol[4 * z + c];
    v = acc / volZ[z];
    phibar[i] = v;
  }
  // A non-finite average becomes NaN and is kept (the residual is then NaN).
  double m = v > 0.0 ? v : 0.0;
  if (!isfinite(v)) m = __longlong_as_double((long long)kNaNBits);
  for (int off = 32; off > 0; off >>= 1) m = nan_keep_max(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0 && (m > 0.0 || m != m))
    atomicMax(maxBits, (unsigned long long)__double_as_longlong(m));
}

__global__ __launch_bounds__(256) void rz_residual_max(const double* __restrict__ phibar,
                                                       double* __restrict__ phibarOld, long n,
                                                       const unsigned long long* __restrict__ maxBits,
                                                       double cutoff,
                                                       unsigned long long* __restrict__ resBits) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (i < n) {
    const double mx = __longlong_as_double((long long)*maxBits);
    const double thr = cutoff * mx;
    const double v = phibar[i];
    if (v > thr) r = fabs((v - phibarOld[i]) / v);
    if (!isfinite(v) || !isfinite(mx)) r = __longlong_as_double((long long)kNaNBits);
    phibarOld[i] = v;
  }
  for (int off = 32; off > 0; off >>= 1) r = nan_keep_max(r, __shfl_down(r, off, 64));
  if ((threadIdx.x & 63) == 0 && (r > 0.0 || r != r))
    atomicMax(resBits, (unsigned long long)__double_as_longlong(r));
}

__global__ void rz_fill_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// ----------------------------------------------------------------------------

static inline unsigned rz_blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename Real>
hipError_t launch_rz_sweep_level(const RzLevelArgs<Real>& a, int K, hipStream_t s) {
  const long threads = 4L * a.nItems * (a.G / K);
  if (threads == 0) return hipSuccess;
  if (threads > kMaxLevelThreads) return hipErrorInvalidValue;
  RzLevelArgs<Real> b = a;
  b.perItem = make_fast_div((unsigned)(a.G / K));
  const dim3 grid(rz_blocks(threads, 256)), block(256);
  switch (K) {
    case 4: hipLaunchKernelGGL((rz_sweep_level<Real, 4>), grid, block, 0, s, b); break;
    case 2: hipLaunchKernelGGL((rz_sweep_level<Real, 2>), grid, block, 0, s, b); break;
    default: hipLaunchKernelGGL((rz_sweep_level<Real, 1>), grid, block, 0, s, b); break;
  }
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_rz_sweep_tile(const RzTileArgs<Real>& a, int nItems, int K, hipStream_t s) {
  if (nItems == 0) return hipSuccess;
  const int blocksPerItem = a.p.G / K;
  if (a.gbPerChunk != 1 && a.gbPerChunk != 2 && a.gbPerChunk != 4) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nItems, (unsigned)((blocksPerItem + a.gbPerChunk - 1) / a.gbPerChunk));
  const dim3 block(64);
  switch (K) {
    case 4: hipLaunchKernelGGL((rz_sweep_tile<Real, 4>), grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL((rz_sweep_tile<Real, 2>), grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL((rz_sweep_tile<Real, 1>), grid, block, 0, s, a); break;
  }
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_rz_source_update(const double* phi, const double* sigs, const double* q, int G,
                                   long n, Real* S, hipStream_t s) {
  hipLaunchKernelGGL((rz_source_update<Real>), dim3(rz_blocks(n, 256)), dim3(256), 0, s, phi,
                     sigs, q, G, n, S);
  return hipGetLastError();
}

hipError_t launch_rz_residual(const double* phi, const double* vol, const double* volZ, int G,
                              long nzg, double* phibar, double* phibarOld, double cutoff,
                              unsigned long long* bits, hipStream_t s) {
  hipLaunchKernelGGL(rz_fill_u64, dim3(1), dim3(64), 0, s, bits, 2);
  hipLaunchKernelGGL(rz_zone_average, dim3(rz_blocks(nzg, 256)), dim3(256), 0, s, phi, vol, volZ,
                     G, nzg, phibar, bits);
  hipLaunchKernelGGL(rz_residual_max, dim3(rz_blocks(nzg, 256)), dim3(256), 0, s, phibar,
                     phibarOld, nzg, bits, cutoff, bits + 1);
  return hipGetLastError();
}

template hipError_t launch_rz_sweep_level<double>(const RzLevelArgs<double>&, int, hipStream_t);
template hipError_t launch_rz_sweep_tile<double>(const RzTileArgs<double>&, int, int, hipStream_t);
template hipError_t launch_rz_sweep_tile<float>(const RzTileArgs<float>&, int, int, hipStream_t);
template hipError_t launch_rz_sweep_level<float>(const RzLevelArgs<float>&, int, hipStream_t);
template hipError_t launch_rz_source_update<double>(const double*, const double*, const double*,
                                                    int, long, double*, hipStream_t);
template hipError_t launch_rz_source_update<float>(const double*, const double*, const double*,
                                                   int, long, float*, hipStream_t);

}  // namespace umt
