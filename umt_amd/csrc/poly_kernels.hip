// The drop-in sweep of one angle on meshes whose zones are not the
// canonical hex: any number of corners per zone (up to kPolyMaxCorner) and
// of corner faces per corner (up to kPolyMaxCF). Teton's general polyhedral
// zones reach the sweep this way; the hex kernels serve the rest.
//
// One thread per (zone, group) of a hyperplane runs the zone as
// snac/SweepUCBxyz.F90:114-305 does: volume sources, incident FP fluxes and
// EZ coupling terms corner by corner, the outgoing EZ face of a corner with
// more than three faces taking the area-weighted mean of the n - 2 incoming
// FP faces opposite it (:196-215), then the corner solve in the schedule's
// corner order (nextC), or from the lagged values of the zone's own corners
// for a zone on the cycle list (:281-295). The arithmetic is the
// reference's, in its order and with no contraction into fused
// multiply-adds, and the quotients are IEEE divisions: the kernel
// reproduces a serial FP64 sweep of the same arrays bit for bit.
//
// The zone's per-corner sources live in per-thread arrays indexed by corner
// number (private memory): this kernel is the general path, not the
// hex-tuned one.
#include "kernels.h"

namespace umt {

__device__ __forceinline__ double poly_dot(double ox, double oy, double oz, const double* a) {
#pragma clang fp contract(off)
  return ox * a[0] + oy * a[1] + oz * a[2];
}

__global__ __launch_bounds__(256) void ucb_sweep_zone_poly(PolySweepArgs p) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)p.nItems * p.G) return;
  const int G = p.G, mcf = p.maxcf;
  const int item = (int)(t / G), g = (int)(t - (long)item * G);
  const int2 it = p.items[item];
  const bool cyclic = it.x < 0;
  const int z = cyclic ? ~it.x : it.x;
  const int nc = p.numCorner[z], c0 = p.cOffSet[z];
  const double ox = p.omega[0], oy = p.omega[1], oz = p.omega[2];
  const double sig = p.sigt[(long)z * G + g];
  constexpr double fouralpha = 1.82;

  double Q[kPolyMaxCorner], src[kPolyMaxCorner], sumArea[kPolyMaxCorner], psiC[kPolyMaxCorner];
  // Volume terms (SweepUCBxyz.F90:116-124).
  for (int c = 0; c < nc; ++c) {
    const long i = (long)(c0 + c) * G + g;
    const double source = p.tauPsi ? p.S[i] + p.tau * p.tauPsi[i] : p.S[i];
    Q[c] = source;
    src[c] = p.vol[c0 + c] * source;
  }

  for (int c = 0; c < nc; ++c) {
    const long cc = c0 + c;
    const int nCF = p.nCFaces[cc];
    double afp[kPolyMaxCF], psifp[kPolyMaxCF];
    double sA = 0.0;
    // Incident fluxes across FP faces (:134-159).
    for (int f = 0; f < nCF; ++f) {
      afp[f] = poly_dot(ox, oy, oz, p.Afp + 3 * (cc * mcf + f));
      if (afp[f] > 0.0) {
        sA = sA + afp[f];
      } else if (afp[f] < 0.0) {
        const int cfp = p.cFP[cc * mcf + f] - 1;
        psifp[f] = p.psi[(long)cfp * G + g];
        src[c] = src[c] - afp[f] * psifp[f];
      }
    }
    // Outgoing EZ faces (:163-252): coupling term to the EZ neighbour.
    for (int f = 0; f < nCF; ++f) {
      const double aez = poly_dot(ox, oy, oz, p.Aez + 3 * (cc * mcf + f));
      if (!(aez > 0.0)) continue;
      const int cez = p.cEZ[cc * mcf + f] - 1;
      sA = sA + aez;
      double areaOpp = 0.0, psiOpp = 0.0;
      if (nCF == 3) {
        const int ifp = (f + 1) % 3;
        if (afp[ifp] < 0.0) {
          psiOpp = psifp[ifp];
          areaOpp = -afp[ifp];
        }
      } else {
        int ifp = f;
        for (int k = 0; k < nCF - 2; ++k) {
          ifp = (ifp + 1) % nCF;
          if (afp[ifp] < 0.0) {
            areaOpp = areaOpp - afp[ifp];
            psiOpp = psiOpp - afp[ifp] * psifp[ifp];
          }
        }
        if (areaOpp > 0.0) psiOpp = psiOpp * (1.0 / areaOpp);
      }
      double sez;
      if (areaOpp > 0.0) {
        const double aez2 = aez * aez;
        const double vol = p.vol[cc];
        const double sigv = sig * vol;
        const double sigv2 = sigv * sigv;
        const double gnum = aez2 * (fouralpha * sigv2 + aez * (4.0 * sigv + 3.0 * aez));
        const double gden = vol * (4.0 * sigv * sigv2 + aez * (6.0 * sigv2 + 2.0 * aez * (2.0 * sigv + aez)));
        sez = (vol * gnum * (sig * psiOpp - Q[c]) + 0.5 * aez * gden * (Q[c] - Q[cez])) /
              (gnum + gden * sig);
      } else {
        const double sigInv = 1.0 / sig;
        sez = 0.5 * aez * sigInv * (Q[c] - Q[cez]);
      }
      src[c] = src[c] + sez;
      src[cez] = src[cez] - sez;
    }
    sumArea[c] = sA;
  }

  // Flux that corner c hands to its downstream EZ neighbours. Each corner
  // pair is coupled through the Omega.A_ez of its lower-numbered corner
  // (:167-179): c -> cez with |Omega.A_ez| when that face is outgoing from c.
  auto handDown = [&](int c, double psi) {
#pragma clang fp contract(off)
    const long cc = c0 + c;
    const int nCF = p.nCFaces[cc];
    for (int f = 0; f < nCF; ++f) {
      const int cez = p.cEZ[cc * mcf + f] - 1;
      double coef = 0.0;
      if (cez > c) {
        const double aez = poly_dot(ox, oy, oz, p.Aez + 3 * (cc * mcf + f));
        if (aez > 0.0) coef = aez;
      } else {
        const long ce = c0 + cez;
        const int nE = p.nCFaces[ce];
        for (int e = 0; e < nE; ++e)
          if (p.cEZ[ce * mcf + e] - 1 == c) {
            const double aez = poly_dot(ox, oy, oz, p.Aez + 3 * (ce * mcf + e));
            if (aez < 0.0) coef = -aez;
          }
      }
      if (coef > 0.0) src[cez] = src[cez] + coef * psi;
    }
  };

  const double SigtOf = sig;
  if (!cyclic) {
    for (int i = 0; i < nc; ++i) {
      const int c = p.nextC[c0 + i] - 1;
      const double psi = src[c] / (sumArea[c] + SigtOf * p.vol[c0 + c]);
      psiC[c] = psi;
      handDown(c, psi);
    }
  } else {
    // Lagged values of the zone's own corners (on the cycle list).
    for (int c = 0; c < nc; ++c) handDown(c, p.psi[(long)(c0 + c) * G + g]);
    for (int c = 0; c < nc; ++c) psiC[c] = src[c] / (sumArea[c] + SigtOf * p.vol[c0 + c]);
  }
  for (int c = 0; c < nc; ++c) {
    const long cc = c0 + c;
    const long i = cc * G + g;
    p.psi[i] = psiC[c];
    p.phi[i] = p.phi[i] + p.quadwt * psiC[c];
    // Exiting boundary fluxes (:298-303) into the boundary tail of psi.
    const int nCF = p.nCFaces[cc];
    for (int f = 0; f < nCF; ++f) {
      const int cfp = p.cFP[cc * mcf + f] - 1;
      if (cfp >= p.C && poly_dot(ox, oy, oz, p.Afp + 3 * (cc * mcf + f)) > 0.0)
        p.psi[(long)cfp * G + g] = psiC[c];
    }
  }
}

hipError_t launch_sweep_poly(const PolySweepArgs& a, hipStream_t s) {
  const long threads = (long)a.nItems * a.G;
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(ucb_sweep_zone_poly, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s,
                     a);
  return hipGetLastError();
}

}  // namespace umt
