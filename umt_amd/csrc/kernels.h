// Kernel argument blocks and launch wrappers for the UCB sweep kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace umt {

template <typename Real>
struct SweepLevelArgs {
  const int2* items;        // (zone, angle) pairs of this level, 0-based
  int nItems;
  int G;                    // groups
  int groupsPerLane;        // K: groups one lane carries (1, 2 or 4; divides G)
  int C;                    // corners; cFP values above C are boundary elements
  const Real* Afp;          // (3, 3, C)
  const Real* Aez;          // (3, 3, C)
  const Real* vol;          // (C)
  const int* cFP;           // (3, C), 1-based contents
  const Real* S;            // (G, C) source per steradian
  const Real* sigt;         // (G, Z)
  const double* omega;      // (3, A)
  Real* psi;                // (G, C + nbelem, A) all angles, or one angle
  long psiAngleStride;      // (C + nbelem) * G
  // Single-angle (drop-in) extras.
  double* phi;              // (G, C), phi += quadwt * psi
  double quadwt;
  const Real* tauPsi;       // (G, C) or null: source += tau * Psi
  Real tau;
};

template <typename Real>
hipError_t launch_sweep_level(const SweepLevelArgs<Real>& a, bool singleAngle, hipStream_t s);
// Nonzero when the sweep kernels can carry K groups per lane for G groups.
int groups_per_lane_supported(int G, int K);
template <typename Real>
hipError_t launch_phi_from_psi(const Real* psi, const double* w, int A, long stride, long n,
                               double* phi, hipStream_t s);
template <typename Real>
hipError_t launch_source_update(const double* phi, const double* sigs, const double* q, int G,
                                long n, Real* S, hipStream_t s);
hipError_t launch_residual(const double* phi, const double* vol, const double* volZ, int G, long nzg,
                           double* phibar, double* phibarOld, double cutoff,
                           unsigned long long* bits, hipStream_t s);
template <typename T>
hipError_t launch_fill_zero(T* p, long n, hipStream_t s);
hipError_t launch_copy_probe(const double* src, double* dst, long n, hipStream_t s);
// phi = sum_a w_a psi (accumulate: phi += ...).
template <typename Real>
hipError_t launch_phi_from_psi_acc(const Real* psi, const double* w, int A, long stride, long n,
                                   double* phi, bool accumulate, hipStream_t s);

// Grey-transport acceleration (gta_kernels.hip).
hipError_t launch_gta_grey_source(const double* phiHalf, const double* phiPrev, const double* sigs,
                                  int G, long n, double* gsrc, hipStream_t s);
template <typename Real>
hipError_t launch_gta_sweep_source(int kind, const double* P, const double* other,
                                   const double* sigs, const double* eps, int G, long n, Real* S,
                                   hipStream_t s);
hipError_t launch_gta_lincomb(const double* x, const double* y, const double* z,
                              const double* coef, int G, long n, double* out, hipStream_t s);
int gta_dot_block(int G);
hipError_t launch_gta_wdot(const double* x, const double* y, const double* sigs, const double* vol,
                           int G, long n, int nblocks, double* partial, double* out, hipStream_t s);
hipError_t launch_gta_zone_error(const double* F, const double* phiHalf, const double* vol,
                                 const double* volZ, int G, long nzg, double* pzOld, int nblocks,
                                 double* partial, double* out, hipStream_t s);
template <typename Dst, typename Src>
hipError_t launch_convert(const Src* s, Dst* d, long n, hipStream_t s_);

}  // namespace umt
