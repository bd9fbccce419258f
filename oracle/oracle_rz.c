/*
 * ORACLE — test infrastructure only (see oracle_rz.h). A line-by-line
 * restatement of the reference's CPU algorithm for the 2D cylindrical (RZ)
 * upstream-corner-balance sweep, written with 1-based index macros so each
 * loop reads like the Fortran it follows. FP64 throughout.
 */
#define _GNU_SOURCE
#include "oracle_rz.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RZ_MAXCF 2
#define RZ_MAXCORNER 4
#define RZ_MAXFACES 4

#define I2(a, i, j, n1) (a)[((i)-1) + (size_t)(n1) * ((j)-1)]
#define I3(a, i, j, k, n1, n2) (a)[((i)-1) + (size_t)(n1) * (((j)-1) + (size_t)(n2) * ((k)-1))]

static double dot2(const double *o, const double *a) { return o[0] * a[0] + o[1] * a[1]; }

/* ------------------------------------------------------------------------ */
/* rt/geometryUCBrz.F90:45-157; zone centre mods/Geometry_mod.F90:344-378 */
int oracle_rz_geometry(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                       const int *cFP, const int *cEZ, const double *px, double *A_fp,
                       double *A_ez, double *RadiusFP, double *RadiusEZ, double *Area,
                       double *Volume) {
  const double half = 0.5, third = 1.0 / 3.0;
  memset(A_fp, 0, sizeof(double) * 4 * (size_t)ncorner);
  memset(A_ez, 0, sizeof(double) * 4 * (size_t)ncorner);
  for (int zone = 1; zone <= nzones; ++zone) {
    const int nCorner = numCorner[zone - 1];
    const int c0 = cOffSet[zone - 1];
    double zc[2] = {0.0, 0.0};
    for (int c = 1; c <= nCorner; ++c) {
      zc[0] += I2(px, 1, c0 + c, 2);
      zc[1] += I2(px, 2, c0 + c, 2);
    }
    zc[0] /= (double)nCorner;
    zc[1] /= (double)nCorner;
    for (int c = 1; c <= nCorner; ++c) {
      const int c1 = I2(cEZ, 1, c0 + c, RZ_MAXCF);
      const int c2 = I2(cEZ, 2, c0 + c, RZ_MAXCF);
      const int cfp1 = I2(cFP, 1, c0 + c, RZ_MAXCF);
      const int cfp2 = I2(cFP, 2, c0 + c, RZ_MAXCF);
      const double r_zone = zc[0], z_zone = zc[1];
      const double r_point = I2(px, 1, c0 + c, 2), z_point = I2(px, 2, c0 + c, 2);
      const double r_point1 = I2(px, 1, c0 + c1, 2), z_point1 = I2(px, 2, c0 + c1, 2);
      const double r_point2 = I2(px, 1, c0 + c2, 2), z_point2 = I2(px, 2, c0 + c2, 2);
      const double r_edge1 = half * (r_point + r_point1), z_edge1 = half * (z_point + z_point1);
      const double r_edge2 = half * (r_point + r_point2), z_edge2 = half * (z_point + z_point2);
      I2(RadiusFP, 1, c0 + c, 2) = half * (r_point + r_edge2);
      I2(RadiusFP, 2, c0 + c, 2) = half * (r_point + r_edge1);
      I2(RadiusEZ, 1, c0 + c, 2) = half * (r_zone + r_edge1);
      I2(RadiusEZ, 2, c0 + c, 2) = half * (r_zone + r_edge2);
      /* the comparisons of a local corner number with a global one are the
       * reference's (geometryUCBrz.F90:102,113); every face is still set */
      if (c < cfp1) {
        I3(A_fp, 1, 1, c0 + c, 2, 2) = half * (z_point2 - z_point);
        I3(A_fp, 2, 1, c0 + c, 2, 2) = half * (r_point - r_point2);
        if (cfp1 <= ncorner) {
          I3(A_fp, 1, 2, cfp1, 2, 2) = -I3(A_fp, 1, 1, c0 + c, 2, 2);
          I3(A_fp, 2, 2, cfp1, 2, 2) = -I3(A_fp, 2, 1, c0 + c, 2, 2);
        }
      }
      if (c < cfp2) {
        I3(A_fp, 1, 2, c0 + c, 2, 2) = half * (z_point - z_point1);
        I3(A_fp, 2, 2, c0 + c, 2, 2) = half * (r_point1 - r_point);
        if (cfp2 <= ncorner) {
          I3(A_fp, 1, 1, cfp2, 2, 2) = -I3(A_fp, 1, 2, c0 + c, 2, 2);
          I3(A_fp, 2, 1, cfp2, 2, 2) = -I3(A_fp, 2, 2, c0 + c, 2, 2);
        }
      }
      if (c < c1) {
        I3(A_ez, 1, 1, c0 + c, 2, 2) = z_edge1 - z_zone;
        I3(A_ez, 2, 1, c0 + c, 2, 2) = r_zone - r_edge1;
        I3(A_ez, 1, 2, c0 + c1, 2, 2) = -I3(A_ez, 1, 1, c0 + c, 2, 2);
        I3(A_ez, 2, 2, c0 + c1, 2, 2) = -I3(A_ez, 2, 1, c0 + c, 2, 2);
      }
      if (c < c2) {
        I3(A_ez, 1, 2, c0 + c, 2, 2) = z_zone - z_edge2;
        I3(A_ez, 2, 2, c0 + c, 2, 2) = r_edge2 - r_zone;
        I3(A_ez, 1, 1, c0 + c2, 2, 2) = -I3(A_ez, 1, 2, c0 + c, 2, 2);
        I3(A_ez, 2, 1, c0 + c2, 2, 2) = -I3(A_ez, 2, 2, c0 + c, 2, 2);
      }
      const double area1 = fabs((r_edge2 - r_point) * (z_zone - z_point) -
                                (z_edge2 - z_point) * (r_zone - r_point));
      const double area2 = fabs((r_zone - r_point) * (z_edge1 - z_point) -
                                (z_zone - z_point) * (r_edge1 - r_point));
      Area[c0 + c - 1] = half * (area1 + area2);
      const double rbar1 = third * (r_point + r_edge2 + r_zone);
      const double rbar2 = third * (r_point + r_edge1 + r_zone);
      Volume[c0 + c - 1] = half * (rbar1 * area1 + rbar2 * area2);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* rt/quadrz.F90:184-268 (product) + rt/rtquad.F90:80-123 (weights to
 * 1/wtiso = 2 pi, Size_mod.F90:278; starting/finishing flags) */
int oracle_rz_quadrature(int npolar, int nazimuthal, const double *cosTheta,
                         const double *weightTheta, const double *cosPhiRZ,
                         const double *weightPhiRZ, double *omega, double *weight, int *starting,
                         int *finishing) {
  if (npolar < 1 || nazimuthal < 1) return 1;
  int m = 0;
  for (int jTheta = npolar; jTheta >= 1; --jTheta) {
    const double cosineTheta = cosTheta[jTheta - 1];
    const double sineTheta = sqrt(1.0 - cosineTheta * cosineTheta);
    const double xilev = cosineTheta;
    for (int s = -1; s <= 1; s += 2) { /* the -xi level, then the +xi level */
      const double xi = s * xilev;
      m += 1;
      I2(omega, 1, m, 2) = -sqrt(1.0 - xilev * xilev);
      I2(omega, 2, m, 2) = xi;
      weight[m - 1] = 0.0;
      for (int iPhi = nazimuthal; iPhi >= 1; --iPhi) {
        m += 1;
        I2(omega, 1, m, 2) = -sineTheta * cosPhiRZ[iPhi - 1];
        I2(omega, 2, m, 2) = xi;
        weight[m - 1] = weightTheta[jTheta - 1] * weightPhiRZ[iPhi - 1];
      }
      for (int iPhi = 1; iPhi <= nazimuthal; ++iPhi) {
        m += 1;
        I2(omega, 1, m, 2) = sineTheta * cosPhiRZ[iPhi - 1];
        I2(omega, 2, m, 2) = xi;
        weight[m - 1] = weightTheta[jTheta - 1] * weightPhiRZ[iPhi - 1];
      }
      m += 1;
      I2(omega, 1, m, 2) = sqrt(1.0 - xilev * xilev);
      I2(omega, 2, m, 2) = xi;
      weight[m - 1] = 0.0;
    }
  }
  const int A = m;
  const double wtiso = 1.0 / (2.0 * M_PI);
  double sumwgt = 0.0;
  for (int ia = 0; ia < A; ++ia) sumwgt += weight[ia];
  const double fac = 1.0 / (wtiso * sumwgt);
  for (int ia = 0; ia < A; ++ia) weight[ia] = fac * weight[ia];
  const double adqtEpsilon = 2.220446049250313e-16; /* epsilon(1.0_adqt) */
  int iang = -1;
  for (int ia = 0; ia < A; ++ia) {
    starting[ia] = 0;
    finishing[ia] = 0;
    if (weight[ia] < adqtEpsilon) {
      if (iang == -1) {
        starting[ia] = 1;
        iang = -iang;
      } else if (iang == 1) {
        finishing[ia] = 1;
        iang = -iang;
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* mods/AngleSet_mod.F90:312-347 with rt/AngleCoef2D.F90:30-100 */
int oracle_rz_angle_coef(int A, const double *omega, const double *weight, const int *starting,
                         const int *finishing, double *angDerivFac, double *quadTauW1,
                         double *quadTauW2) {
  double *Alpha = calloc(A, sizeof(double)), *Tau = calloc(A, sizeof(double));
  int rc = 0, totalAngles = 0;
  while (totalAngles < A) {
    /* one xi level: from its starting direction up to the next one */
    int nlevelAngles = 1;
    while (totalAngles + nlevelAngles < A && !starting[totalAngles + nlevelAngles]) ++nlevelAngles;
    const int angle1 = totalAngles + 1, angle2 = totalAngles + nlevelAngles;
    double weightLevel = 0.0, Phimh = M_PI, Mumh = I2(omega, 1, 1, 2);
    for (int angle = angle1; angle <= angle2; ++angle) weightLevel += weight[angle - 1];
    for (int angle = angle1; angle <= angle2; ++angle) {
      if (starting[angle - 1]) {
        Alpha[angle - 1] = 0.0;
        Tau[angle - 1] = 0.0;
        if (angle == angle1) {
          Phimh = M_PI;
          Mumh = I2(omega, 1, angle, 2);
        } else {
          rc = 2; /* "Mu not increasing in xi-level" */
          goto out;
        }
      } else if (finishing[angle - 1]) {
        Alpha[angle - 1] = 0.0;
        Tau[angle - 1] = 0.0;
      } else {
        Alpha[angle - 1] = Alpha[angle - 2] - weight[angle - 1] * I2(omega, 1, angle, 2);
        const double Phiph = Phimh - weight[angle - 1] * M_PI / weightLevel;
        const double xi = I2(omega, 2, angle, 2);
        const double Muph = sqrt(1.0 - xi * xi) * cos(Phiph);
        const double mu = I2(omega, 1, angle, 2);
        if (mu < Mumh || mu > Muph) {
          rc = 3; /* "Mu not between limits" */
          goto out;
        }
        Tau[angle - 1] = (mu - Mumh) / (Muph - Mumh);
        Phimh = Phiph;
        Mumh = Muph;
      }
    }
    totalAngles += nlevelAngles;
  }
  for (int n = 0; n < A; ++n) {
    if (starting[n] || finishing[n]) {
      angDerivFac[n] = 0.0;
      quadTauW1[n] = 1.0;
      quadTauW2[n] = 0.0;
    } else {
      angDerivFac[n] = I2(omega, 1, n + 1, 2) + Alpha[n] / (weight[n] * Tau[n]);
      quadTauW1[n] = 1.0 / Tau[n];
      quadTauW2[n] = (1.0 - Tau[n]) / Tau[n];
    }
  }
out:
  free(Alpha);
  free(Tau);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* snac/snneed.F90:62-93 (2D), snac/findseeds.F90:58-66,
 * snac/getDownStreamData.F90:88-149 (nCFaces = 2), snac/snnext.F90:103-195 */
int oracle_rz_schedule_angle(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                             const int *zoneFaces, const int *cEZ, const int *CToFace,
                             const int *zoneOpp, const int *faceOpp, const double *A_fp,
                             const double *A_ez, const double *omega, int *nextZ, int *nextC,
                             int *zonesInPlane, int *nHyperPlanesOut) {
  (void)ncorner;
  int *needZ = calloc(nzones, sizeof(int));
  int *listZone = calloc(nzones, sizeof(int));
  char *exitFace = calloc((size_t)RZ_MAXFACES * nzones, 1);
  char *doneZ = calloc(nzones, 1);
  int rc = 0;
  for (int zone = 1; zone <= nzones; ++zone) {
    const int nCorner = numCorner[zone - 1], c0 = cOffSet[zone - 1];
    for (int c = 1; c <= nCorner; ++c) {
      const int face = I2(CToFace, 1, c0 + c, RZ_MAXCF);
      const int zOpp = I2(zoneOpp, face, zone, RZ_MAXFACES);
      if (zone < zOpp) {
        const int fOpp = I2(faceOpp, face, zone, RZ_MAXFACES);
        const double afpm = dot2(omega, &I3(A_fp, 1, 1, c0 + c, 2, 2));
        if (afpm < 0.0) {
          needZ[zone - 1] += 1;
          I2(exitFace, fOpp, zOpp, RZ_MAXFACES) = 1;
        } else if (afpm > 0.0) {
          needZ[zOpp - 1] += 1;
          I2(exitFace, face, zone, RZ_MAXFACES) = 1;
        }
      }
    }
  }
  int newZones = 0;
  for (int zone = 1; zone <= nzones; ++zone)
    if (needZ[zone - 1] == 0) listZone[newZones++] = zone;
  if (newZones == 0) { rc = 3; goto done; }
  for (int zone = 1; zone <= nzones; ++zone) {
    int need[RZ_MAXCORNER] = {0}, nDSC[RZ_MAXCORNER] = {0}, DownStreamC[RZ_MAXCORNER][RZ_MAXCF];
    const int nCorner = numCorner[zone - 1], c0 = cOffSet[zone - 1];
    for (int c = 1; c <= nCorner; ++c) {
      for (int cface = 1; cface <= 2; ++cface) {
        const int cez = I2(cEZ, cface, c0 + c, RZ_MAXCF);
        if (cez > c) {
          const double aez = dot2(omega, &I3(A_ez, 1, cface, c0 + c, 2, 2));
          if (aez < 0.0) {
            need[c - 1] += 1;
            DownStreamC[cez - 1][nDSC[cez - 1]++] = c;
          } else if (aez > 0.0) {
            need[cez - 1] += 1;
            DownStreamC[c - 1][nDSC[c - 1]++] = cez;
          }
        }
      }
    }
    int badZone = 0;
    for (int i = 1; i <= nCorner; ++i) {
      int c = 1;
      for (int k = 2; k <= nCorner; ++k)
        if (need[k - 1] < need[c - 1]) c = k;
      if (need[c - 1] != 0) badZone = 1;
      nextC[c0 + i - 1] = c;
      for (int cf = 0; cf < nDSC[c - 1]; ++cf) need[DownStreamC[c - 1][cf] - 1] -= 1;
      need[c - 1] = 99;
    }
    if (badZone) { rc = 4; goto done; }
  }
  {
    int ndoneZ = 0, nextZone = 0, lastZone = 0, nHyperPlanes = 0;
    for (;;) {
      nHyperPlanes += 1;
      zonesInPlane[nHyperPlanes - 1] = newZones;
      nextZone = lastZone + newZones;
      int addedZones = 0;
      for (int zID = 1; zID <= newZones; ++zID) {
        const int zone = listZone[lastZone + zID - 1];
        ndoneZ += 1;
        doneZ[zone - 1] = 1;
        for (int face = 1; face <= zoneFaces[zone - 1]; ++face) {
          if (I2(exitFace, face, zone, RZ_MAXFACES)) {
            const int Zexit = I2(zoneOpp, face, zone, RZ_MAXFACES);
            if (Zexit > 0 && !doneZ[Zexit - 1]) {
              needZ[Zexit - 1] -= 1;
              if (needZ[Zexit - 1] == 0) {
                nextZone += 1;
                addedZones += 1;
                listZone[nextZone - 1] = Zexit;
              } else if (needZ[Zexit - 1] < 0) {
                rc = 5;
                goto done;
              }
            }
          }
        }
        nextZ[ndoneZ - 1] = zone;
      }
      lastZone += newZones;
      if (lastZone == nzones) break;
      if (addedZones == 0) { rc = 6; goto done; }
      newZones = addedZones;
    }
    *nHyperPlanesOut = nHyperPlanes;
  }
done:
  free(needZ); free(listZone); free(exitFace); free(doneZ);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* snac/SweepUCBrz.F90:105-276 for one angle. Exiting boundary fluxes are
 * taken from the corner faces with Omega.A_fp > 0 on the boundary (the
 * BdyExit list of rt/findexit.F90:299-345). */
void oracle_rz_sweep_angle(int nzones, int ncorner, int nbelem, int Groups, const int *numCorner,
                           const int *cOffSet, const int *cFP, const int *cEZ, const double *A_fp,
                           const double *A_ez, const double *RadiusFP, const double *RadiusEZ,
                           const double *Area, const double *Volume, const double *STotal,
                           const double *Sigt, double tau, const double *Psi,
                           const double *omega, double quadwt, double fac, double quadTauW1,
                           double quadTauW2, int startingDirection, int nHyperPlanes,
                           const int *zonesInPlane, const int *nextZ, const int *nextC,
                           double *Psi1, double *PsiM, double *PsiB, double *PsiBnext,
                           double *Phi, double *PsiNext) {
  (void)nzones;
  const double fouralpha = 1.82;
  const int G = Groups;
  double *Q = malloc(sizeof(double) * G * RZ_MAXCORNER);
  double *src = malloc(sizeof(double) * G * RZ_MAXCORNER);
  double sumArea[RZ_MAXCORNER], coefpsi[2][RZ_MAXCORNER];
  int nxez[RZ_MAXCORNER], ez_exit[2][RZ_MAXCORNER];

  for (int ib = 1; ib <= nbelem; ++ib)
    for (int g = 1; g <= G; ++g) I2(Psi1, g, ncorner + ib, G) = I2(PsiB, g, ib, G);

  int ndoneZ = 0;
  for (int hyperPlane = 1; hyperPlane <= nHyperPlanes; ++hyperPlane) {
    const int nz = zonesInPlane[hyperPlane - 1];
    for (int ii = 1; ii <= nz; ++ii) {
      const int zone = abs(nextZ[ndoneZ + ii - 1]);
      const int nCorner = numCorner[zone - 1];
      const int c0 = cOffSet[zone - 1];
      for (int c = 0; c < RZ_MAXCORNER; ++c) nxez[c] = 0;
      for (int c = 1; c <= nCorner; ++c) {
        for (int g = 1; g <= G; ++g) {
          double source = I2(STotal, g, c0 + c, G);
          if (tau != 0.0) source += tau * I2(Psi, g, c0 + c, G);
          I2(Q, g, c, G) = source;
          I2(src, g, c, G) = Volume[c0 + c - 1] * source;
        }
        sumArea[c - 1] = fac * Area[c0 + c - 1];
      }
      for (int c = 1; c <= nCorner; ++c) {
        for (int cface = 1; cface <= 2; ++cface) {
          const double afp = dot2(omega, &I3(A_fp, 1, cface, c0 + c, 2, 2));
          const double aez = dot2(omega, &I3(A_ez, 1, cface, c0 + c, 2, 2));
          int cfp = -1;
          if (afp < 0.0) {
            cfp = I2(cFP, cface, c0 + c, RZ_MAXCF);
            const double R_afp = I2(RadiusFP, cface, c0 + c, 2) * afp;
            sumArea[c - 1] = sumArea[c - 1] - R_afp;
            for (int g = 1; g <= G; ++g)
              I2(src, g, c, G) = I2(src, g, c, G) - R_afp * I2(Psi1, g, cfp, G);
          }
          if (aez > 0.0) {
            const double R = I2(RadiusEZ, cface, c0 + c, 2);
            const int cez = I2(cEZ, cface, c0 + c, RZ_MAXCF);
            const double area = Area[c0 + c - 1];
            nxez[c - 1] += 1;
            ez_exit[nxez[c - 1] - 1][c - 1] = cez;
            coefpsi[nxez[c - 1] - 1][c - 1] = R * aez;
            sumArea[cez - 1] = sumArea[cez - 1] + R * aez;
            if (afp < 0.0) {
              for (int g = 1; g <= G; ++g) {
                const double sig = I2(Sigt, g, zone, G);
                const double sigA = sig * area;
                const double sigA2 = sigA * sigA;
                const double gnum = aez * aez * (fouralpha * sigA2 + aez * (4.0 * sigA + 3.0 * aez));
                const double gden =
                    area * (4.0 * sigA * sigA2 + aez * (6.0 * sigA2 + 2.0 * aez * (2.0 * sigA + aez)));
                const double sez = R * (area * gnum * (sig * I2(Psi1, g, cfp, G) - I2(Q, g, c, G)) +
                                        0.5 * aez * gden * (I2(Q, g, c, G) - I2(Q, g, cez, G))) /
                                   (gnum + gden * sig);
                I2(src, g, c, G) = I2(src, g, c, G) + sez;
                I2(src, g, cez, G) = I2(src, g, cez, G) - sez;
              }
            } else {
              for (int g = 1; g <= G; ++g) {
                const double sez =
                    0.5 * R * aez * (I2(Q, g, c, G) - I2(Q, g, cez, G)) / I2(Sigt, g, zone, G);
                I2(src, g, c, G) = I2(src, g, c, G) + sez;
                I2(src, g, cez, G) = I2(src, g, cez, G) - sez;
              }
            }
          }
        }
      }
      for (int i = 1; i <= nCorner; ++i) {
        const int c = nextC[c0 + i - 1];
        for (int g = 1; g <= G; ++g)
          I2(Psi1, g, c0 + c, G) =
              (I2(src, g, c, G) + Area[c0 + c - 1] * fac * I2(PsiM, g, c0 + c, G)) /
              (sumArea[c - 1] + I2(Sigt, g, zone, G) * Volume[c0 + c - 1]);
        for (int cface = 1; cface <= nxez[c - 1]; ++cface) {
          const int cez = ez_exit[cface - 1][c - 1];
          for (int g = 1; g <= G; ++g)
            I2(src, g, cez, G) = I2(src, g, cez, G) + coefpsi[cface - 1][c - 1] * I2(Psi1, g, c0 + c, G);
        }
      }
      if (startingDirection) {
        for (int c = 1; c <= nCorner; ++c)
          for (int g = 1; g <= G; ++g) I2(PsiM, g, c0 + c, G) = I2(Psi1, g, c0 + c, G);
      } else {
        for (int c = 1; c <= nCorner; ++c)
          for (int g = 1; g <= G; ++g) {
            I2(PsiM, g, c0 + c, G) =
                quadTauW1 * I2(Psi1, g, c0 + c, G) - quadTauW2 * I2(PsiM, g, c0 + c, G);
            I2(Phi, g, c0 + c, G) = I2(Phi, g, c0 + c, G) + quadwt * I2(Psi1, g, c0 + c, G);
          }
      }
    }
    ndoneZ += nz;
  }
  /* exiting boundary fluxes; for a finishing next direction its PsiB is PsiM */
  for (int c = 1; c <= ncorner; ++c)
    for (int cface = 1; cface <= 2; ++cface) {
      const int cfp = I2(cFP, cface, c, RZ_MAXCF);
      if (cfp > ncorner && dot2(omega, &I3(A_fp, 1, cface, c, 2, 2)) > 0.0) {
        const int b = cfp - ncorner;
        for (int g = 1; g <= G; ++g) I2(PsiB, g, b, G) = I2(Psi1, g, c, G);
        if (PsiBnext)
          for (int g = 1; g <= G; ++g) I2(PsiBnext, g, b, G) = I2(PsiM, g, c, G);
      }
    }
  if (PsiNext) /* savePsi with a finishing next direction: Psi(:,:,Angle+1) = PsiM */
    memcpy(PsiNext, PsiM, sizeof(double) * (size_t)G * ncorner);
  free(Q);
  free(src);
}

/* ------------------------------------------------------------------------ */
/* snac/SetSweep.F90:90-170 for one RZ domain with vacuum boundaries: PsiM
 * and Phi start at zero, the angles of each xi level are swept in order
 * (finishing directions are not swept), xi levels shared among threads
 * with a Phi per thread summed in thread order. With PsiPrev (G, C, A) not
 * NULL, each direction's source is STotal + tau PsiPrev(:, :, Angle)
 * (SweepUCBrz.F90:136; Sigt is then the caller's sigma_t + tau). Psi (G, C, A) receives every
 * angle's psi when not NULL (finishing directions: the last PsiM). */
void oracle_rz_sweep_all(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                         const int *cOffSet, const int *cFP, const int *cEZ, const double *A_fp,
                         const double *A_ez, const double *RadiusFP, const double *RadiusEZ,
                         const double *Area, const double *Volume, const double *STotal,
                         const double *Sigt, const double *omega, const double *weight,
                         const int *starting, const int *finishing, const double *angDerivFac,
                         const double *quadTauW1, const double *quadTauW2,
                         const int *nHyperPlanes, const int *zonesInPlane, const int *nextZ,
                         const int *nextC, double tau, const double *PsiPrev, int nthreads,
                         double *PsiB, double *Phi, double *Psi) {
  const size_t nGC = (size_t)G * ncorner, nPsi1 = (size_t)G * (ncorner + nbelem);
  int *planeStart = malloc(sizeof(int) * (A + 1));
  planeStart[0] = 0;
  for (int a = 0; a < A; ++a) planeStart[a + 1] = planeStart[a] + nHyperPlanes[a];
  int nLevels = 0;
  int *levelStart = malloc(sizeof(int) * (A + 1));
  for (int a = 0; a < A; ++a)
    if (starting[a]) levelStart[nLevels++] = a;
  levelStart[nLevels] = A;
  if (nthreads < 1) nthreads = 1;
  double *phiSets = calloc((size_t)nthreads * nGC, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    double *Psi1 = calloc(nPsi1, sizeof(double));
    double *PsiM = calloc(nGC, sizeof(double));
    double *phiT = phiSets + (size_t)t * nGC;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int l = 0; l < nLevels; ++l) {
      memset(PsiM, 0, sizeof(double) * nGC);
      for (int a = levelStart[l]; a < levelStart[l + 1]; ++a) {
        if (finishing[a]) continue;
        const int finNext = a + 1 < A && finishing[a + 1];
        oracle_rz_sweep_angle(nzones, ncorner, nbelem, G, numCorner, cOffSet, cFP, cEZ, A_fp,
                              A_ez, RadiusFP, RadiusEZ, Area, Volume, STotal, Sigt,
                              PsiPrev ? tau : 0.0, PsiPrev ? &PsiPrev[(size_t)a * nGC] : NULL,
                              &omega[2 * (size_t)a], weight[a], angDerivFac[a], quadTauW1[a],
                              quadTauW2[a], starting[a], nHyperPlanes[a],
                              &zonesInPlane[planeStart[a]], &nextZ[(size_t)a * nzones],
                              &nextC[(size_t)a * ncorner], Psi1, PsiM,
                              &PsiB[(size_t)a * G * nbelem],
                              finNext ? &PsiB[(size_t)(a + 1) * G * nbelem] : NULL, phiT,
                              (Psi && finNext) ? &Psi[(size_t)(a + 1) * nGC] : NULL);
        if (Psi) memcpy(&Psi[(size_t)a * nGC], Psi1, sizeof(double) * nGC);
      }
    }
    free(Psi1);
    free(PsiM);
  }
  for (size_t i = 0; i < nGC; ++i) {
    double s = 0.0;
    for (int t = 0; t < nthreads; ++t) s += phiSets[(size_t)t * nGC + i];
    Phi[i] = s;
  }
  free(phiSets);
  free(planeStart);
  free(levelStart);
}

/* ------------------------------------------------------------------------ */
/* Steady source iteration as oracle_source_iteration, in RZ: STotal =
 * wtiso (sigs phi + q) with wtiso = 1/(2 pi) (mods/Size_mod.F90:277-279),
 * residual over volume-weighted zone averages of the four corners. */
int oracle_rz_source_iteration(int nzones, int ncorner, int nbelem, int G, int A,
                               const int *numCorner, const int *cOffSet, const int *cFP,
                               const int *cEZ, const double *A_fp, const double *A_ez,
                               const double *RadiusFP, const double *RadiusEZ, const double *Area,
                               const double *Volume, const double *VolumeZone, const double *Sigt,
                               const double *sigs, const double *q, const double *omega,
                               const double *weight, const int *starting, const int *finishing,
                               const double *angDerivFac, const double *quadTauW1,
                               const double *quadTauW2, const int *nHyperPlanes,
                               const int *zonesInPlane, const int *nextZ, const int *nextC,
                               int nthreads, double tol, int maxIters, double *Phi,
                               double *residual, double *hist) {
  const size_t nGC = (size_t)G * ncorner, nGZ = (size_t)G * nzones;
  const double wtiso = 1.0 / (2.0 * M_PI);
  double *STotal = malloc(sizeof(double) * nGC);
  double *PsiB = calloc((size_t)G * nbelem * A, sizeof(double));
  double *old = malloc(sizeof(double) * nGZ), *cur = malloc(sizeof(double) * nGZ);
  for (int z = 1; z <= nzones; ++z)
    for (int g = 1; g <= G; ++g) {
      double s = 0.0;
      for (int c = 1; c <= numCorner[z - 1]; ++c)
        s += I2(Phi, g, cOffSet[z - 1] + c, G) * Volume[cOffSet[z - 1] + c - 1];
      I2(old, g, z, G) = s / VolumeZone[z - 1];
    }
  int it = 0;
  double res = 1.0;
  while (it < maxIters) {
    for (int z = 1; z <= nzones; ++z)
      for (int c = 1; c <= numCorner[z - 1]; ++c)
        for (int g = 1; g <= G; ++g)
          I2(STotal, g, cOffSet[z - 1] + c, G) =
              wtiso * (I2(sigs, g, z, G) * I2(Phi, g, cOffSet[z - 1] + c, G) + I2(q, g, z, G));
    memset(PsiB, 0, sizeof(double) * (size_t)G * nbelem * A);
    oracle_rz_sweep_all(nzones, ncorner, nbelem, G, A, numCorner, cOffSet, cFP, cEZ, A_fp, A_ez,
                        RadiusFP, RadiusEZ, Area, Volume, STotal, Sigt, omega, weight, starting,
                        finishing, angDerivFac, quadTauW1, quadTauW2, nHyperPlanes, zonesInPlane,
                        nextZ, nextC, 0.0, NULL, nthreads, PsiB, Phi, NULL);
    double mx = 0.0;
    for (int z = 1; z <= nzones; ++z)
      for (int g = 1; g <= G; ++g) {
        double s = 0.0;
        for (int c = 1; c <= numCorner[z - 1]; ++c)
          s += I2(Phi, g, cOffSet[z - 1] + c, G) * Volume[cOffSet[z - 1] + c - 1];
        I2(cur, g, z, G) = s / VolumeZone[z - 1];
        if (I2(cur, g, z, G) > mx) mx = I2(cur, g, z, G);
      }
    const double thr = 1e-6 * mx;
    res = 0.0;
    for (size_t i = 0; i < nGZ; ++i) {
      if (cur[i] > thr) {
        const double r = fabs((cur[i] - old[i]) / cur[i]);
        if (r > res) res = r;
      }
      old[i] = cur[i];
    }
    if (hist) hist[it] = res;
    ++it;
    if (res < tol) break;
  }
  if (residual) *residual = res;
  free(STotal); free(PsiB); free(old); free(cur);
  return it;
}
