// Kernels of the grey-transport acceleration (GTA) solve, applied per group:
// each group is its own one-group ("grey") problem, for which Teton's grey
// opacities reduce to that group's sigma_t and sigma_s
// (rt/setGTAOpacity.F90:195-233 with one group and no emission coupling).
//
// Vectors are (G, C) FP64 arrays, groups fastest. Inner products are the
// sigma_s V weighted sums of rt/scat_prod.F90 / rt/scat_prod1.F90, one per
// group, reduced in a fixed order so every run gives the same bits.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace umt {

namespace {
constexpr double kWtiso = 0.07957747154594767;  // 1/(4 pi)

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// max that keeps a NaN once one is seen (the host checks for it).
__device__ inline double nan_max(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return fmax(a, b);
}
}  // namespace

// Scattering-rate residual after a sweep: sigma_s (phi_half - phi_prev)
// (rt/getCollisionRate.F90:64,87 for one group).
__global__ __launch_bounds__(256) void gta_grey_source(const double* __restrict__ phiHalf,
                                                       const double* __restrict__ phiPrev,
                                                       const double* __restrict__ sigs, int G,
                                                       long n, double* __restrict__ gsrc) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c = i / G;
  const int g = (int)(i - c * G);
  gsrc[i] = sigs[(c >> 3) * G + g] * (phiHalf[i] - phiPrev[i]);
}

// Sources of the two grey sweeps (snac/GTASweep.F90:89-102):
//   kind 1: wtiso (sigma_s P + gsrc)          (gsrc may be null)
//   kind 2: wtiso sigma_s eps (P1 - Pold)     (stretched TSA step)
template <typename Real>
__global__ __launch_bounds__(256) void gta_sweep_source(int kind, const double* __restrict__ P,
                                                        const double* __restrict__ other,
                                                        const double* __restrict__ sigs,
                                                        const double* __restrict__ eps, int G,
                                                        long n, Real* __restrict__ S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c = i / G;
  const int g = (int)(i - c * G);
  const long zg = (c >> 3) * G + g;
  const double p = P ? P[i] : 0.0;  // null P stands for zero
  double v;
  if (kind == 1)
    v = kWtiso * (sigs[zg] * p + (other ? other[i] : 0.0));
  else
    v = kWtiso * sigs[zg] * eps[zg] * (other[i] - p);
  S[i] = (Real)v;
}

// out = a_g x + b_g y + c_g z with per-group coefficients coef[3g..3g+2];
// null inputs count as zero. out may alias any input.
__global__ __launch_bounds__(256) void gta_lincomb(const double* x, const double* y,
                                                   const double* z,
                                                   const double* __restrict__ coef, int G,
                                                   long n, double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = (int)(i % G);
  double v = 0.0;
  if (x) v += coef[3 * g + 0] * x[i];
  if (y) v += coef[3 * g + 1] * y[i];
  if (z) v += coef[3 * g + 2] * z[i];
  out[i] = v;
}

// Per-group sums of x * y * sigma_s V (y null: x * sigma_s V). Block size is
// a multiple of G and the grid stride a multiple of the block size, so each
// thread always sees the same group; partial[block][g] is summed in thread
// order, then gta_reduce_partials sums the blocks in block order.
__global__ void gta_wdot_partial(const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ sigs, const double* __restrict__ vol,
                                 int G, long n, double* __restrict__ partial) {
  extern __shared__ double red[];
  double acc = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = i / G;
    const int g = (int)(i - c * G);
    const double w = sigs[(c >> 3) * G + g] * vol[c];
    acc += x[i] * (y ? y[i] : 1.0) * w;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if ((int)threadIdx.x < G) {
    double s = 0.0;
    for (int t = threadIdx.x; t < (int)blockDim.x; t += G) s += red[t];
    partial[(long)blockIdx.x * G + threadIdx.x] = s;
  }
}

// Both of BiCGStab's omega inner products in one pass: per-group sums of
// x y sigma_s V (partial[0 .. nblocks G)) and x x sigma_s V (partial[nblocks G
// .. 2 nblocks G)), each reduced like gta_wdot_partial's.
__global__ void gta_wdot2_partial(const double* __restrict__ x, const double* __restrict__ y,
                                  const double* __restrict__ sigs, const double* __restrict__ vol,
                                  int G, long n, double* __restrict__ partial) {
  extern __shared__ double red[];
  double axy = 0.0, axx = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = i / G;
    const int g = (int)(i - c * G);
    const double w = sigs[(c >> 3) * G + g] * vol[c];
    const double xi = x[i];
    axy += xi * y[i] * w;
    axx += xi * xi * w;
  }
  red[threadIdx.x] = axy;
  red[blockDim.x + threadIdx.x] = axx;
  __syncthreads();
  if ((int)threadIdx.x < G) {
    double s1 = 0.0, s2 = 0.0;
    for (int t = threadIdx.x; t < (int)blockDim.x; t += G) {
      s1 += red[t];
      s2 += red[blockDim.x + t];
    }
    partial[(long)blockIdx.x * G + threadIdx.x] = s1;
    partial[((long)gridDim.x + blockIdx.x) * G + threadIdx.x] = s2;
  }
}

// BiCGStab's closing updates of one step in one pass, per group g with
// coef[3g] = alpha, coef[3g+1] = omega: F += alpha d + omega r, then
// r -= omega As, with the per-group sums of the new r sigma_s V reduced
// like gta_wdot_partial's (the next step's r.r-hat product).
__global__ void gta_bicg_update(double* __restrict__ F, const double* __restrict__ d,
                                double* __restrict__ r, const double* __restrict__ As,
                                const double* __restrict__ coef, const double* __restrict__ sigs,
                                const double* __restrict__ vol, int G, long n,
                                double* __restrict__ partial) {
  extern __shared__ double red[];
  double acc = 0.0;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = i / G;
    const int g = (int)(i - c * G);
    const double alpha = coef[3 * g], omega = coef[3 * g + 1];
    const double ri = r[i];
    F[i] = F[i] + alpha * d[i] + omega * ri;
    const double rn = ri - omega * As[i];
    r[i] = rn;
    acc += rn * (sigs[(c >> 3) * G + g] * vol[c]);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if ((int)threadIdx.x < G) {
    double s = 0.0;
    for (int t = threadIdx.x; t < (int)blockDim.x; t += G) s += red[t];
    partial[(long)blockIdx.x * G + threadIdx.x] = s;
  }
}

// One block per group: each thread sums a fixed stride of the block
// partials, then a fixed-shape tree in LDS, so the result has the same bits
// on every run.
__global__ __launch_bounds__(256) void gta_reduce_partials(const double* __restrict__ partial,
                                                           int nblocks, int G,
                                                           double* __restrict__ out) {
  __shared__ double red[256];
  const int g = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partial[(long)b * G + g];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[g] = red[0];
}

// Grey convergence measure (rt/GTASolver.F90:318-356 per group): zone
// correction pz, change against the previous pz, relative to the zone
// average of phi_half + pz; per-block max of the pointwise relative change
// and sums of V errZone^2 and V phiNew^2. pzOld is updated.
__global__ void gta_zone_error(const double* __restrict__ F, const double* __restrict__ phiHalf,
                               const double* __restrict__ vol, const double* __restrict__ volZ,
                               int G, long nzg, double* __restrict__ pzOld,
                               double* __restrict__ partial /* [block][3] */) {
  __shared__ double smax[256], serr[256], sphi[256];
  double mx = 0.0, e2 = 0.0, p2 = 0.0;
  bool bad = false;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nzg; i += stride) {
    const long z = i / G;
    const int g = (int)(i - z * G);
    double pz = 0.0, rad = 0.0;
    for (int c = 0; c < 8; ++c) {
      const long k = (8 * z + c) * G + g;
      pz += vol[8 * z + c] * F[k];
      rad += vol[8 * z + c] * phiHalf[k];
    }
    pz /= volZ[z];
    rad /= volZ[z];
    const double err = pz - pzOld[i];
    const double phiNew = rad + pz;
    e2 += volZ[z] * err * err;
    p2 += volZ[z] * phiNew * phiNew;
    // A NaN or Inf anywhere poisons the maximum: the host checks isfinite.
    if (!isfinite(err) || !isfinite(phiNew)) bad = true;
    else if (phiNew != 0.0) mx = fmax(mx, fabs(err / phiNew));
    pzOld[i] = pz;
  }
  smax[threadIdx.x] = bad ? __builtin_nan("") : mx;
  serr[threadIdx.x] = e2;
  sphi[threadIdx.x] = p2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0, a = 0.0, b = 0.0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      const double v = smax[t];
      m = nan_max(m, v);
      a += serr[t];
      b += sphi[t];
    }
    partial[3 * blockIdx.x + 0] = m;
    partial[3 * blockIdx.x + 1] = a;
    partial[3 * blockIdx.x + 2] = b;
  }
}

__global__ __launch_bounds__(256) void gta_reduce_error(const double* __restrict__ partial,
                                                        int nblocks, double* __restrict__ out) {
  __shared__ double sm[256], sa[256], sb[256];
  double m = 0.0, a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nblocks; k += blockDim.x) {
    m = nan_max(m, partial[3 * k]);
    a += partial[3 * k + 1];
    b += partial[3 * k + 2];
  }
  sm[threadIdx.x] = m;
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sm[threadIdx.x] = nan_max(sm[threadIdx.x], sm[threadIdx.x + w]);
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sm[0];
    out[1] = sa[0];
    out[2] = sb[0];
  }
}

// ----------------------------------------------------------------------------
// Transfer-matrix solver (the reference's "new" GTA: per-zone corner transfer
// matrices, snac/InitSweepGreyUCBxyz.F90:10-253, and each zone's direct
// solve with in-place LU factors, snac/UpdateScalarIntensity.F90:11-170).
//
// T(s -> c) of a zone and group is the scalar flux at corner c from a unit
// per-steradian source at corner s of the same zone with no incident flux,
// summed over the grey directions. Column s of every zone comes from one
// sweep of the grey directions with "no inflow" (every zone on its own, so
// all (zone, direction) items go in one launch) of the unit source at local
// corner s. T is stored as 8 arrays (G, C): T(s -> c) of zone z, group g at
// s * C * G + (8 z + c) * G + g; the LU factors replace T element by element.

// S = 1 at local corner s of every zone and group, 0 elsewhere.
template <typename Real>
__global__ __launch_bounds__(256) void gta_unit_source(int s, int G, long n, Real* __restrict__ S) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long c = i / G;
  S[i] = (Real)((c & 7) == s ? 1 : 0);
}

// A = I - wtiso sigma_s T per (zone, group), A(c, s) = delta_cs - wtiso
// sigma_s T(s -> c) (the zone's sigma_s is the same at all its corners),
// factored in place as A = L U (Doolittle, no pivoting: unit lower L below
// the diagonal, U on and above it), in the reference's loop order
// (UpdateScalarIntensity.F90:64-97). Element (c, s) lives where T(s -> c)
// was. One thread per (zone, group).
__global__ __launch_bounds__(128) void gta_zone_lu(double* __restrict__ T,
                                                   const double* __restrict__ sigs, int G,
                                                   long nzg, long CG) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nzg) return;
  const long z = t / G;
  const int g = (int)(t - z * G);
  double* base = T + (8 * z) * G + g;
  const double ws = kWtiso * sigs[t];
  double a[8][8];
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int s = 0; s < 8; ++s) a[c][s] = (c == s ? 1.0 : 0.0) - ws * base[s * CG + (long)c * G];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    double u = 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) u += a[i][k] * a[k][i];
    a[i][i] -= u;
    const double dinv = 1.0 / a[i][i];
#pragma unroll
    for (int j = i + 1; j < 8; ++j) {
      double r = 0.0, v = 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) {
        r += a[i][k] * a[k][j];
        v += a[j][k] * a[k][i];
      }
      a[i][j] -= r;
      a[j][i] = dinv * (a[j][i] - v);
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int s = 0; s < 8; ++s) base[s * CG + (long)c * G] = a[c][s];
}

// out = P + (L U)^-1 (phi - P) per (zone, group): the new solver's grey
// operator (P null stands for zero). phi is the grey sweep's scalar flux with
// source wtiso (sigma_s P [+ gsrc]); P + (I - B)^-1 (phi - P), B = wtiso
// T sigma_s, equals the reference's (I - B)^-1 (PhiInc [+ T wtiso gsrc])
// because the sweep is linear: PhiInc = phi - T wtiso (sigma_s P [+ gsrc]).
// Forward and back substitution as ScalarIntensitySolve
// (UpdateScalarIntensity.F90:136-167). out may alias phi. residual != 0
// writes P - M(P) = -(I - B)^-1 (phi - P) instead, the operator's residual
// form that BiCGStab applies (v - M v).
__global__ __launch_bounds__(256) void gta_zone_solve(const double* __restrict__ LU,
                                                      const double* P, const double* phi, int G,
                                                      long nzg, long CG, double* out,
                                                      int residual) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nzg) return;
  const long z = t / G;
  const int g = (int)(t - z * G);
  const double* base = LU + (8 * z) * G + g;
  const long k0 = (8 * z) * G + g;
  double p[8], x[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    p[c] = P ? P[k0 + (long)c * G] : 0.0;
    x[c] = phi[k0 + (long)c * G] - p[c];
  }
#pragma unroll
  for (int j = 1; j < 8; ++j) {
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < j; ++i) r -= base[i * CG + (long)j * G] * x[i];
    x[j] += r;
  }
  x[7] = x[7] / base[7 * CG + 7L * G];
#pragma unroll
  for (int k = 6; k >= 0; --k) {
    double r = 0.0;
#pragma unroll
    for (int i = k + 1; i < 8; ++i) r += x[i] * base[i * CG + (long)k * G];
    x[k] = (x[k] - r) / base[k * CG + (long)k * G];
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) out[k0 + (long)c * G] = residual ? -x[c] : p[c] + x[c];
}

// ----------------------------------------------------------------------------

hipError_t launch_gta_grey_source(const double* phiHalf, const double* phiPrev, const double* sigs,
                                  int G, long n, double* gsrc, hipStream_t s) {
  hipLaunchKernelGGL(gta_grey_source, dim3(blocks_for(n, 256)), dim3(256), 0, s, phiHalf, phiPrev,
                     sigs, G, n, gsrc);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_gta_sweep_source(int kind, const double* P, const double* other,
                                   const double* sigs, const double* eps, int G, long n, Real* S,
                                   hipStream_t s) {
  hipLaunchKernelGGL((gta_sweep_source<Real>), dim3(blocks_for(n, 256)), dim3(256), 0, s, kind, P,
                     other, sigs, eps, G, n, S);
  return hipGetLastError();
}

hipError_t launch_gta_lincomb(const double* x, const double* y, const double* z,
                              const double* coef, int G, long n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(gta_lincomb, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, y, z, coef, G, n,
                     out);
  return hipGetLastError();
}

int gta_dot_block(int G) { return G * (256 / G > 0 ? 256 / G : 1); }

hipError_t launch_gta_wdot(const double* x, const double* y, const double* sigs, const double* vol,
                           int G, long n, int nblocks, double* partial, double* out,
                           hipStream_t s) {
  const int bs = gta_dot_block(G);
  hipLaunchKernelGGL(gta_wdot_partial, dim3(nblocks), dim3(bs), bs * sizeof(double), s, x, y, sigs,
                     vol, G, n, partial);
  hipLaunchKernelGGL(gta_reduce_partials, dim3(G), dim3(256), 0, s, partial, nblocks, G, out);
  return hipGetLastError();
}

hipError_t launch_gta_wdot2(const double* x, const double* y, const double* sigs, const double* vol,
                            int G, long n, int nblocks, double* partial, double* out, hipStream_t s) {
  const int bs = gta_dot_block(G);
  hipLaunchKernelGGL(gta_wdot2_partial, dim3(nblocks), dim3(bs), 2 * bs * sizeof(double), s, x, y,
                     sigs, vol, G, n, partial);
  hipLaunchKernelGGL(gta_reduce_partials, dim3(G), dim3(256), 0, s, partial, nblocks, G, out);
  hipLaunchKernelGGL(gta_reduce_partials, dim3(G), dim3(256), 0, s, partial + (long)nblocks * G,
                     nblocks, G, out + G);
  return hipGetLastError();
}

hipError_t launch_gta_bicg_update(double* F, const double* d, double* r, const double* As,
                                  const double* coef, const double* sigs, const double* vol, int G,
                                  long n, int nblocks, double* partial, double* out, hipStream_t s) {
  const int bs = gta_dot_block(G);
  hipLaunchKernelGGL(gta_bicg_update, dim3(nblocks), dim3(bs), bs * sizeof(double), s, F, d, r, As,
                     coef, sigs, vol, G, n, partial);
  hipLaunchKernelGGL(gta_reduce_partials, dim3(G), dim3(256), 0, s, partial, nblocks, G, out);
  return hipGetLastError();
}

hipError_t launch_gta_zone_error(const double* F, const double* phiHalf, const double* vol,
                                 const double* volZ, int G, long nzg, double* pzOld, int nblocks,
                                 double* partial, double* out, hipStream_t s) {
  hipLaunchKernelGGL(gta_zone_error, dim3(nblocks), dim3(256), 0, s, F, phiHalf, vol, volZ, G, nzg,
                     pzOld, partial);
  hipLaunchKernelGGL(gta_reduce_error, dim3(1), dim3(256), 0, s, partial, nblocks, out);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_gta_unit_source(int s, int G, long n, Real* S, hipStream_t st) {
  hipLaunchKernelGGL((gta_unit_source<Real>), dim3(blocks_for(n, 256)), dim3(256), 0, st, s, G, n, S);
  return hipGetLastError();
}

hipError_t launch_gta_zone_lu(double* T, const double* sigs, int G, long nzg, long CG,
                              hipStream_t s) {
  hipLaunchKernelGGL(gta_zone_lu, dim3(blocks_for(nzg, 128)), dim3(128), 0, s, T, sigs, G, nzg, CG);
  return hipGetLastError();
}

hipError_t launch_gta_zone_solve(const double* LU, const double* P, const double* phi, int G,
                                 long nzg, long CG, double* out, bool residual, hipStream_t s) {
  hipLaunchKernelGGL(gta_zone_solve, dim3(blocks_for(nzg, 256)), dim3(256), 0, s, LU, P, phi, G,
                     nzg, CG, out, residual ? 1 : 0);
  return hipGetLastError();
}

template hipError_t launch_gta_unit_source<double>(int, int, long, double*, hipStream_t);
template hipError_t launch_gta_unit_source<float>(int, int, long, float*, hipStream_t);

template hipError_t launch_gta_sweep_source<double>(int, const double*, const double*,
                                                    const double*, const double*, int, long,
                                                    double*, hipStream_t);
template hipError_t launch_gta_sweep_source<float>(int, const double*, const double*, const double*,
                                                   const double*, int, long, float*, hipStream_t);

}  // namespace umt
