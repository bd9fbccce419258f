/*
 * ORACLE — test infrastructure only (see oracle.h). A line-by-line
 * restatement of the reference's CPU algorithm for the UCB xyz sweep path,
 * written with 1-based index macros so each loop reads like the Fortran it
 * follows. FP64 throughout; no FMA contraction is relied on.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Leading dimensions of the corner-face, corner and zone-face arrays
 * (Teton's maxcf, maxCorner, maxFaces). The hex build takes 3, 8, 6; the
 * polyhedral build (liboracle_poly.so) is compiled with larger values. */
#ifndef MAXCF
#define MAXCF 3
#endif
#ifndef MAXCORNER
#define MAXCORNER 8
#endif
#ifndef MAXFACES
#define MAXFACES 6
#endif

/* 1-based column-major accessors */
#define I2(a, i, j, n1) (a)[((i)-1) + (size_t)(n1) * ((j)-1)]
#define I3(a, i, j, k, n1, n2) (a)[((i)-1) + (size_t)(n1) * (((j)-1) + (size_t)(n2) * ((k)-1))]

static double dot3(const double *o, const double *a) { return o[0] * a[0] + o[1] * a[1] + o[2] * a[2]; }

/* Omega.A, zero when |Omega.A| <= tol |A|_1 (face parallel to Omega). */
static double dot3_tol(const double *o, const double *a, double tol) {
  const double d = dot3(o, a);
  if (tol > 0.0 && fabs(d) <= tol * (fabs(a[0]) + fabs(a[1]) + fabs(a[2]))) return 0.0;
  return d;
}

/* ------------------------------------------------------------------------ */
/* rt/geometryUCBxyz.F90:63-181; centres from mods/Geometry_mod.F90:344-441 */
int oracle_geometry(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                    const int *zoneFaces, const int *nCFaces, const int *cFP, const int *cEZ,
                    const int *CToFace, const int *zoneOpp, const double *px, double *A_fp,
                    double *A_ez, double *Volume) {
  memset(A_fp, 0, sizeof(double) * 3 * MAXCF * (size_t)ncorner);
  for (int zone = 1; zone <= nzones; ++zone) {
    const int nFaces = zoneFaces[zone - 1];
    const int nCorner = numCorner[zone - 1];
    const int c0 = cOffSet[zone - 1];
    double zoneCenter[3] = {0, 0, 0};
    double faceCenter[MAXFACES][3];
    int ncf[MAXFACES];
    for (int c = 1; c <= nCorner; ++c)
      for (int d = 0; d < 3; ++d) zoneCenter[d] += I2(px, d + 1, c0 + c, 3);
    for (int d = 0; d < 3; ++d) zoneCenter[d] /= (double)nCorner;
    memset(faceCenter, 0, sizeof faceCenter);
    memset(ncf, 0, sizeof ncf);
    for (int c = 1; c <= nCorner; ++c)
      for (int cf = 1; cf <= nCFaces[c0 + c - 1]; ++cf) {
        const int face = I2(CToFace, cf, c0 + c, MAXCF);
        ncf[face - 1] += 1;
        for (int d = 0; d < 3; ++d) faceCenter[face - 1][d] += I2(px, d + 1, c0 + c, 3);
      }
    for (int f = 0; f < nFaces; ++f)
      for (int d = 0; d < 3; ++d) faceCenter[f][d] = ncf[f] ? faceCenter[f][d] / (double)ncf[f] : 0.0;

    for (int c = 1; c <= nCorner; ++c) {
      Volume[c0 + c - 1] = 0.0;
      for (int cf = 1; cf <= MAXCF; ++cf)
        for (int d = 1; d <= 3; ++d) I3(A_ez, d, cf, c0 + c, 3, MAXCF) = 0.0;
    }
    for (int c = 1; c <= nCorner; ++c) {
      const int cc = c0 + c;
      const int nCF = nCFaces[cc - 1];
      for (int cface1 = 1; cface1 <= nCF; ++cface1) {
        const int cface2 = cface1 % nCF + 1;
        const int cfp = I2(cFP, cface1, cc, MAXCF);
        const int cez1 = I2(cEZ, cface1, cc, MAXCF);
        const int cez2 = I2(cEZ, cface2, cc, MAXCF);
        const int face = I2(CToFace, cface1, cc, MAXCF);
        double pnt0[3], pnt1[3], pnt2[3], tdl[3], tfl[3], tzl[3], tfz[3], tfe1[3], tfe2[3];
        double A_fep[3], A_fez1[3], A_fez2[3];
        for (int d = 0; d < 3; ++d) {
          pnt0[d] = I2(px, d + 1, c0 + c, 3);
          pnt1[d] = I2(px, d + 1, c0 + cez1, 3);
          pnt2[d] = I2(px, d + 1, c0 + cez2, 3);
        }
        for (int d = 0; d < 3; ++d) {
          tdl[d] = 0.5 * (pnt1[d] - pnt2[d]);
          tfl[d] = faceCenter[face - 1][d] - pnt0[d];
          tzl[d] = zoneCenter[d] - pnt0[d];
          tfz[d] = faceCenter[face - 1][d] - zoneCenter[d];
          tfe1[d] = faceCenter[face - 1][d] - 0.5 * (pnt1[d] + pnt0[d]);
          tfe2[d] = faceCenter[face - 1][d] - 0.5 * (pnt2[d] + pnt0[d]);
        }
        A_fep[0] = 0.5 * (tfl[2] * tdl[1] - tfl[1] * tdl[2]);
        A_fep[1] = 0.5 * (tfl[0] * tdl[2] - tfl[2] * tdl[0]);
        A_fep[2] = 0.5 * (tfl[1] * tdl[0] - tfl[0] * tdl[1]);
        const int zOpp = I2(zoneOpp, face, zone, MAXFACES);
        if (zOpp > 0) {
          if (zOpp > zone) {
            for (int d = 1; d <= 3; ++d) I3(A_fp, d, cface1, cc, 3, MAXCF) = A_fep[d - 1];
            for (int cf = 1; cf <= nCFaces[cfp - 1]; ++cf)
              if (I2(cFP, cf, cfp, MAXCF) == cc)
                for (int d = 1; d <= 3; ++d) I3(A_fp, d, cf, cfp, 3, MAXCF) = -A_fep[d - 1];
          } else {
            for (int d = 1; d <= 3; ++d) A_fep[d - 1] = I3(A_fp, d, cface1, cc, 3, MAXCF);
          }
        } else if (zOpp < 0) {
          for (int d = 1; d <= 3; ++d) I3(A_fp, d, cface1, cc, 3, MAXCF) = A_fep[d - 1];
        }
        A_fez1[0] = 0.5 * (tfz[2] * tfe1[1] - tfz[1] * tfe1[2]);
        A_fez1[1] = 0.5 * (tfz[0] * tfe1[2] - tfz[2] * tfe1[0]);
        A_fez1[2] = 0.5 * (tfz[1] * tfe1[0] - tfz[0] * tfe1[1]);
        A_fez2[0] = 0.5 * (tfz[1] * tfe2[2] - tfz[2] * tfe2[1]);
        A_fez2[1] = 0.5 * (tfz[2] * tfe2[0] - tfz[0] * tfe2[2]);
        A_fez2[2] = 0.5 * (tfz[0] * tfe2[1] - tfz[1] * tfe2[0]);
        for (int d = 1; d <= 3; ++d) {
          I3(A_ez, d, cface1, cc, 3, MAXCF) += A_fez1[d - 1];
          I3(A_ez, d, cface2, cc, 3, MAXCF) += A_fez2[d - 1];
        }
        Volume[cc - 1] += (1.0 / 3.0) * fabs(tzl[0] * A_fep[0] + tzl[1] * A_fep[1] + tzl[2] * A_fep[2]);
      }
    }
    for (int c = 1; c <= nCorner; ++c) {
      const int cc = c0 + c;
      for (int cface1 = 1; cface1 <= nCFaces[cc - 1]; ++cface1) {
        const int cez = I2(cEZ, cface1, cc, MAXCF);
        if (cez > c)
          for (int cface2 = 1; cface2 <= nCFaces[c0 + cez - 1]; ++cface2)
            if (I2(cEZ, cface2, c0 + cez, MAXCF) == c)
              for (int d = 1; d <= 3; ++d)
                I3(A_ez, d, cface2, c0 + cez, 3, MAXCF) = -I3(A_ez, d, cface1, cc, 3, MAXCF);
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* rt/quadProduct.F90:52-182, rt/rtquad.F90:95-105 */
int oracle_quadrature(int npolar, int nazimuthal, int polaraxis, const double *cosTheta,
                      const double *weightTheta, const double *cosPhi, const double *weightPhi,
                      double *omega, double *weight) {
  if (npolar < 1 || nazimuthal < 1 || polaraxis < 1 || polaraxis > 3) return 1;
  const int nangoct = npolar * nazimuthal;
  double *ox = malloc(sizeof(double) * nangoct), *oy = malloc(sizeof(double) * nangoct);
  double *oz = malloc(sizeof(double) * nangoct), *qw = malloc(sizeof(double) * nangoct);
  int m = 0;
  for (int iPhi = 1; iPhi <= nazimuthal; ++iPhi) {
    const double cosinePhi = cosPhi[iPhi - 1];
    for (int jTheta = npolar; jTheta >= 1; --jTheta) {
      const double cosineTheta = cosTheta[jTheta - 1];
      const double sineTheta = sqrt(1.0 - cosineTheta * cosineTheta);
      if (polaraxis == 1) {
        ox[m] = cosineTheta;
        oy[m] = sineTheta * cosinePhi;
        oz[m] = sqrt(1.0 - ox[m] * ox[m] - oy[m] * oy[m]);
      } else if (polaraxis == 2) {
        oy[m] = cosineTheta;
        oz[m] = sineTheta * cosinePhi;
        ox[m] = sqrt(1.0 - oy[m] * oy[m] - oz[m] * oz[m]);
      } else {
        oz[m] = cosineTheta;
        ox[m] = sineTheta * cosinePhi;
        oy[m] = sqrt(1.0 - ox[m] * ox[m] - oz[m] * oz[m]);
      }
      qw[m] = weightTheta[jTheta - 1] * weightPhi[iPhi - 1];
      ++m;
    }
  }
  static const int sx[8] = {1, -1, -1, 1, 1, -1, -1, 1};
  static const int sy[8] = {1, 1, -1, -1, 1, 1, -1, -1};
  static const int sz[8] = {1, 1, 1, 1, -1, -1, -1, -1};
  int nn = 0;
  for (int i = 0; i < nangoct; ++i) {
    for (int o = 0; o < 8; ++o) {
      I2(omega, 1, nn + o + 1, 3) = sx[o] * ox[i];
      I2(omega, 2, nn + o + 1, 3) = sy[o] * oy[i];
      I2(omega, 3, nn + o + 1, 3) = sz[o] * oz[i];
      weight[nn + o] = qw[i];
    }
    nn += 8;
  }
  const int nang = 8 * nangoct;
  double sumwgt = 0.0;
  for (int ia = 0; ia < nang; ++ia) sumwgt += weight[ia];
  const double wtiso = 1.0 / (4.0 * M_PI);
  const double fac = 1.0 / (wtiso * sumwgt);
  for (int ia = 0; ia < nang; ++ia) weight[ia] = fac * weight[ia];
  free(ox); free(oy); free(oz); free(qw);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* snneed (3D) + findseeds + getDownStreamData + snnext, one angle */
int oracle_schedule_angle(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                          const int *zoneFaces, const int *nCFaces, const int *cEZ,
                          const int *CToFace, const int *zoneOpp, const int *faceOpp,
                          const double *A_fp, const double *A_ez, const double *omega,
                          int *nextZ, int *nextC, int *zonesInPlane, int *nHyperPlanesOut) {
  return oracle_schedule_angle_tol(nzones, ncorner, numCorner, cOffSet, zoneFaces, nCFaces, cEZ,
                                   CToFace, zoneOpp, faceOpp, A_fp, A_ez, omega, 0.0, nextZ,
                                   nextC, zonesInPlane, nHyperPlanesOut);
}

int oracle_schedule_angle_tol(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                              const int *zoneFaces, const int *nCFaces, const int *cEZ,
                              const int *CToFace, const int *zoneOpp, const int *faceOpp,
                              const double *A_fp, const double *A_ez, const double *omega,
                              double parTol, int *nextZ, int *nextC, int *zonesInPlane,
                              int *nHyperPlanesOut) {
  int *cycleList = malloc(sizeof(int) * (size_t)ncorner);
  int numCycles = 0;
  int rc = oracle_schedule_angle_cyc(nzones, ncorner, numCorner, cOffSet, zoneFaces, nCFaces, cEZ,
                                     CToFace, zoneOpp, faceOpp, A_fp, A_ez, omega, parTol, nextZ,
                                     nextC, zonesInPlane, nHyperPlanesOut, cycleList, &numCycles);
  free(cycleList);
  if (rc == 0 && numCycles > 0) rc = 2;
  return rc;
}

/* State shared by snnext and the routines it calls (their argument lists). */
typedef struct {
  int nzones, ncorner;
  const int *numCorner, *cOffSet, *zoneFaces, *zoneOpp, *faceOpp;
  int *needZ;
  char *exitFace;    /* exitFace(MAXFACES, nzones) */
  char *onCycleList; /* (nzones) */
  int *cycleList;    /* 1-based corners, capacity cycleCap */
  int meshCycles, cycleCap, overflow;
  /* cyclebreaker / sccsearch work arrays */
  int *dfnum, *lowlink, *stack, *tempList, *zoneBreakList;
  char *isNew, *onstack;
  int ncount, stackindex, nBreaks;
} SchedState;

/* The corners of one zone onto the cycle list (snac/fixZone.F90:24-30 and the
 * same loop in snneed, findseeds and sccsearch). */
static void add_zone_corners(SchedState *st, int zone) {
  for (int c = 1; c <= st->numCorner[zone - 1]; ++c) {
    if (st->meshCycles >= st->cycleCap) { st->overflow = 1; return; }
    st->cycleList[st->meshCycles++] = st->cOffSet[zone - 1] + c;
  }
}

/* snac/sccsearch.F90:54-183 (Tarjan's strongly connected components; on a
 * component of more than one zone, break the links into its root). */
static void sccsearch(SchedState *st, int zone) {
  st->ncount += 1;
  st->dfnum[zone - 1] = st->ncount;
  st->lowlink[zone - 1] = st->ncount;
  st->isNew[zone - 1] = 0;
  st->stackindex += 1;
  st->stack[st->stackindex - 1] = zone;
  st->onstack[zone - 1] = 1;

  const int nFaces = st->zoneFaces[zone - 1];
  for (int face = 1; face <= nFaces; ++face) { /* :96-134 */
    if (I2(st->exitFace, face, zone, MAXFACES)) {
      const int zone2 = I2(st->zoneOpp, face, zone, MAXFACES);
      if (zone2 > 0) {
        if (st->isNew[zone2 - 1]) {
          sccsearch(st, zone2);
          if (st->lowlink[zone2 - 1] < st->lowlink[zone - 1])
            st->lowlink[zone - 1] = st->lowlink[zone2 - 1];
        } else {
          if (st->dfnum[zone2 - 1] < st->dfnum[zone - 1] && st->onstack[zone2 - 1] &&
              st->lowlink[zone2 - 1] < st->lowlink[zone - 1])
            st->lowlink[zone - 1] = st->lowlink[zone2 - 1];
        }
      }
    }
  }

  if (st->lowlink[zone - 1] == st->dfnum[zone - 1]) { /* :138-181 */
    int zone2 = st->stack[st->stackindex - 1];
    st->stackindex -= 1;
    st->onstack[zone2 - 1] = 0;
    if (zone2 != zone) {
      int cyclesize = 0;
      while (zone2 != zone) {
        cyclesize += 1;
        st->tempList[cyclesize - 1] = zone2;
        zone2 = st->stack[st->stackindex - 1];
        st->stackindex -= 1;
      }
      cyclesize += 1;
      st->tempList[cyclesize - 1] = zone2;
      st->onstack[st->tempList[0] - 1] = 1;

      const int lowlinkZ = st->tempList[cyclesize - 1];
      const int nF = st->zoneFaces[lowlinkZ - 1];
      for (int face = 1; face <= nF; ++face) {
        const int zoneBreak = I2(st->zoneOpp, face, lowlinkZ, MAXFACES);
        const int faceBreak = I2(st->faceOpp, face, lowlinkZ, MAXFACES);
        if (zoneBreak > 0) {
          if (st->onstack[zoneBreak - 1] && I2(st->exitFace, faceBreak, zoneBreak, MAXFACES)) {
            if (!st->onCycleList[zoneBreak - 1]) {
              add_zone_corners(st, zoneBreak);
              st->onCycleList[zoneBreak - 1] = 1;
            }
            st->needZ[lowlinkZ - 1] -= 1;
            I2(st->exitFace, faceBreak, zoneBreak, MAXFACES) = 0;
            if (st->needZ[lowlinkZ - 1] == 0) {
              st->nBreaks += 1;
              st->zoneBreakList[st->nBreaks - 1] = lowlinkZ;
            }
          }
        }
      }
      for (int i = 1; i <= cyclesize; ++i) st->onstack[st->tempList[i - 1] - 1] = 0;
    }
  }
}

/* snac/cyclebreaker.F90:50-150. Returns addedZones (> 0) or a negative error. */
static int cyclebreaker(SchedState *st, int ndoneZ, int *nextZone, int *listZone) {
  const int nzones = st->nzones;
  const int ngraph = nzones - ndoneZ;
  int *listZ = malloc(sizeof(int) * (size_t)(ngraph > 0 ? ngraph : 1));
  for (int z = 0; z < nzones; ++z) {
    st->isNew[z] = 1;
    st->onstack[z] = 0;
    st->dfnum[z] = 0;
    st->lowlink[z] = 0;
  }
  st->nBreaks = 0;
  st->ncount = 0;
  st->stackindex = 0;
  int nleft = 0;
  for (int zone = 1; zone <= nzones; ++zone) {
    if (st->needZ[zone - 1] == 0) st->isNew[zone - 1] = 0;
    else {
      if (nleft < ngraph) listZ[nleft] = zone;
      nleft += 1;
    }
  }
  int rc = 0;
  if (nleft != ngraph) { rc = -1; goto out; }
  for (int i = 1; i <= ngraph; ++i) {
    const int zone = listZ[i - 1];
    if (st->isNew[zone - 1]) sccsearch(st, zone);
  }
  if (st->nBreaks == 0) { rc = -2; goto out; }
  int addedZones = 0;
  for (int i = 1; i <= st->nBreaks; ++i) {
    const int zone = st->zoneBreakList[i - 1];
    if (st->needZ[zone - 1] == 0) {
      *nextZone += 1;
      addedZones += 1;
      listZone[*nextZone - 1] = zone;
    } else if (st->needZ[zone - 1] < 0) {
      rc = -3;
      goto out;
    }
  }
  rc = addedZones == 0 ? -4 : addedZones;
out:
  free(listZ);
  return rc;
}

int oracle_schedule_angle_cyc(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                              const int *zoneFaces, const int *nCFaces, const int *cEZ,
                              const int *CToFace, const int *zoneOpp, const int *faceOpp,
                              const double *A_fp, const double *A_ez, const double *omega,
                              double parTol, int *nextZ, int *nextC, int *zonesInPlane,
                              int *nHyperPlanesOut, int *cycleList, int *numCyclesOut) {
  SchedState st = {0};
  st.nzones = nzones;
  st.ncorner = ncorner;
  st.numCorner = numCorner;
  st.cOffSet = cOffSet;
  st.zoneFaces = zoneFaces;
  st.zoneOpp = zoneOpp;
  st.faceOpp = faceOpp;
  st.needZ = calloc(nzones, sizeof(int));
  st.exitFace = calloc((size_t)MAXFACES * nzones, 1);
  st.onCycleList = calloc(nzones, 1);
  /* fixZone appends without the onCycleList test, so a zone can appear twice */
  st.cycleCap = 2 * ncorner + MAXFACES * MAXCORNER;
  st.cycleList = malloc(sizeof(int) * (size_t)st.cycleCap);
  st.dfnum = calloc(nzones, sizeof(int));
  st.lowlink = calloc(nzones, sizeof(int));
  st.stack = calloc(nzones, sizeof(int));
  st.tempList = calloc(nzones, sizeof(int));
  st.zoneBreakList = calloc(nzones, sizeof(int));
  st.isNew = calloc(nzones, 1);
  st.onstack = calloc(nzones, 1);
  int *listZone = calloc(nzones, sizeof(int));
  char *badZone = calloc(nzones, 1);
  char *doneZ = calloc(nzones, 1);
  int rc = 0;

  /* snac/snneed.F90:97-191 */
  for (int zone = 1; zone <= nzones; ++zone) {
    const int nCorner = numCorner[zone - 1], c0 = cOffSet[zone - 1], nFaces = zoneFaces[zone - 1];
    double afpm_Face[MAXFACES] = {0};
    int nInc[MAXFACES] = {0}, nExit[MAXFACES] = {0};
    for (int c = 1; c <= nCorner; ++c) {
      const int cc = c0 + c;
      for (int cface = 1; cface <= nCFaces[cc - 1]; ++cface) {
        const int face = I2(CToFace, cface, cc, MAXCF);
        const int zOpp = I2(zoneOpp, face, zone, MAXFACES);
        if (zOpp > zone) {
          const double afpm = dot3_tol(omega, &I3(A_fp, 1, cface, cc, 3, MAXCF), parTol);
          afpm_Face[face - 1] += afpm;
          if (afpm < 0.0) nInc[face - 1] += 1;
          else if (afpm > 0.0) nExit[face - 1] += 1;
        }
      }
    }
    for (int face = 1; face <= nFaces; ++face) {
      const int zOpp = I2(zoneOpp, face, zone, MAXFACES);
      if (zOpp > zone) {
        const int fOpp = I2(faceOpp, face, zone, MAXFACES);
        if (afpm_Face[face - 1] < 0.0) {
          st.needZ[zone - 1] += 1;
          I2(st.exitFace, fOpp, zOpp, MAXFACES) = 1;
          if (nExit[face - 1] > 0 && !st.onCycleList[zone - 1]) {
            add_zone_corners(&st, zone);
            st.onCycleList[zone - 1] = 1;
          }
        } else if (afpm_Face[face - 1] > 0.0) {
          st.needZ[zOpp - 1] += 1;
          I2(st.exitFace, face, zone, MAXFACES) = 1;
          if (nInc[face - 1] > 0 && !st.onCycleList[zOpp - 1]) {
            add_zone_corners(&st, zOpp);
            st.onCycleList[zOpp - 1] = 1;
          }
        }
      }
    }
  }

  /* snac/findseeds.F90:58-114 */
  int newZones = 0;
  for (int zone = 1; zone <= nzones; ++zone)
    if (st.needZ[zone - 1] == 0) listZone[newZones++] = zone;
  if (newZones == 0) {
    int minNeed = nzones, zoneID = 0;
    for (int zone = 1; zone <= nzones; ++zone) {
      int boundaryZone = 0; /* aux/setOppositeFace.F90:41-45 */
      for (int face = 1; face <= zoneFaces[zone - 1]; ++face)
        if (I2(zoneOpp, face, zone, MAXFACES) < 0) boundaryZone = 1;
      if (boundaryZone && st.needZ[zone - 1] < minNeed) {
        zoneID = zone;
        minNeed = st.needZ[zone - 1];
      }
    }
    if (zoneID == 0) { rc = 3; goto done; }
    newZones = 1;
    listZone[0] = zoneID;
    st.needZ[zoneID - 1] = 0;
    for (int face = 1; face <= zoneFaces[zoneID - 1]; ++face) {
      if (!I2(st.exitFace, face, zoneID, MAXFACES)) {
        const int zOpp = I2(zoneOpp, face, zoneID, MAXFACES);
        const int fOpp = I2(faceOpp, face, zoneID, MAXFACES);
        if (zOpp > 0) {
          add_zone_corners(&st, zOpp);
          I2(st.exitFace, fOpp, zOpp, MAXFACES) = 0;
          st.onCycleList[zOpp - 1] = 1;
        }
      }
    }
  }

  /* snac/getDownStreamData.F90:88-158 */
  for (int zone = 1; zone <= nzones; ++zone) {
    int need[MAXCORNER] = {0}, nDSC[MAXCORNER] = {0}, DownStreamC[MAXCORNER][MAXCF];
    const int nCorner = numCorner[zone - 1], c0 = cOffSet[zone - 1];
    for (int c = 1; c <= nCorner; ++c) {
      const int cc = c0 + c;
      for (int cface = 1; cface <= nCFaces[cc - 1]; ++cface) {
        const int cez = I2(cEZ, cface, cc, MAXCF);
        if (cez > c) {
          const double aez = dot3_tol(omega, &I3(A_ez, 1, cface, cc, 3, MAXCF), parTol);
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
    for (int i = 1; i <= nCorner; ++i) {
      int c = 1; /* minloc: first index of the minimum */
      for (int k = 2; k <= nCorner; ++k)
        if (need[k - 1] < need[c - 1]) c = k;
      if (need[c - 1] != 0) badZone[zone - 1] = 1;
      nextC[c0 + i - 1] = c;
      for (int cf = 0; cf < nDSC[c - 1]; ++cf) need[DownStreamC[c - 1][cf] - 1] -= 1;
      need[c - 1] = 99;
    }
    if (badZone[zone - 1]) {
      for (int i = 1; i <= nCorner; ++i) nextC[c0 + i - 1] = i;
      add_zone_corners(&st, zone); /* snac/fixZone.F90 */
    }
  }

  /* snac/snnext.F90:103-213 */
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
          if (I2(st.exitFace, face, zone, MAXFACES)) {
            const int Zexit = I2(zoneOpp, face, zone, MAXFACES);
            if (Zexit > 0 && !doneZ[Zexit - 1]) {
              st.needZ[Zexit - 1] -= 1;
              if (st.needZ[Zexit - 1] == 0) {
                nextZone += 1;
                addedZones += 1;
                listZone[nextZone - 1] = Zexit;
              } else if (st.needZ[Zexit - 1] < 0) {
                rc = 5;
                goto done;
              }
            }
          }
        }
        nextZ[ndoneZ - 1] = badZone[zone - 1] ? -zone : zone;
      }
      lastZone += newZones;
      if (lastZone == nzones) break;
      if (addedZones > 0) {
        newZones = addedZones;
      } else {
        addedZones = cyclebreaker(&st, ndoneZ, &nextZone, listZone);
        if (addedZones < 0) { rc = 6; goto done; }
        newZones = addedZones;
      }
    }
    *nHyperPlanesOut = nHyperPlanes;
  }
  if (st.overflow || st.meshCycles > ncorner) { rc = 7; goto done; }
  memcpy(cycleList, st.cycleList, sizeof(int) * (size_t)st.meshCycles);
  *numCyclesOut = st.meshCycles;
done:
  free(st.needZ); free(st.exitFace); free(st.onCycleList); free(st.cycleList);
  free(st.dfnum); free(st.lowlink); free(st.stack); free(st.tempList); free(st.zoneBreakList);
  free(st.isNew); free(st.onstack);
  free(listZone); free(badZone); free(doneZ);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* snac/SweepUCBxyz.F90:82-318 */
static void sweep_angle_tol(int nzones, int ncorner, int nbelem, int Groups, const int *numCorner,
                            const int *cOffSet, const int *nCFacesArray, const int *cFP,
                            const int *cEZ, const double *A_fp, const double *A_ez,
                            const double *Volume, const double *STotal, const double *Sigt,
                            double tau, const double *Psi, const double *omega, double quadwt,
                            int nHyperPlanes, const int *zonesInPlane, const int *nextZ,
                            const int *nextC, double parTol, int numCycles, const int *cycleList,
                            double *cyclePsi, double *Psi1, double *PsiB, double *Phi) {
  (void)nzones;
  const double fouralpha = 1.82;
  const int G = Groups;
  double *psi_opp = calloc(G, sizeof(double));
  double *SigtVol = malloc(sizeof(double) * G * MAXCORNER);
  double *src = malloc(sizeof(double) * G * MAXCORNER);
  double *Q = malloc(sizeof(double) * G * MAXCORNER);
  double *psifp = calloc((size_t)G * MAXCF, sizeof(double));
  double sumArea[MAXCORNER], afp[MAXCF], coefpsi[MAXCF][MAXCORNER];
  int nxez[MAXCORNER], ez_exit[MAXCF][MAXCORNER], bdy_exit[2][MAXCF * MAXCORNER];

  for (int ib = 1; ib <= nbelem; ++ib)
    for (int g = 1; g <= G; ++g) I2(Psi1, g, ncorner + ib, G) = I2(PsiB, g, ib, G);
  /* initFromCycleList (control/constructDynMemory.F90:118-152): corners read
   * before they are computed hold the previous sweep's values. */
  for (int m = 1; m <= numCycles; ++m)
    for (int g = 1; g <= G; ++g) I2(Psi1, g, cycleList[m - 1], G) = I2(cyclePsi, g, m, G);

  int ndoneZ = 0;
  for (int hyperPlane = 1; hyperPlane <= nHyperPlanes; ++hyperPlane) {
    const int nz = zonesInPlane[hyperPlane - 1];
    for (int ii = 1; ii <= nz; ++ii) {
      const int zone0 = nextZ[ndoneZ + ii - 1];
      const int zone = abs(zone0);
      const int nCorner = numCorner[zone - 1];
      const int c0 = cOffSet[zone - 1];
      int nxBdy = 0;
      for (int c = 1; c <= nCorner; ++c)
        for (int g = 1; g <= G; ++g) {
          double source = I2(STotal, g, c0 + c, G);
          if (tau != 0.0) source += tau * I2(Psi, g, c0 + c, G);
          I2(Q, g, c, G) = source;
          I2(src, g, c, G) = Volume[c0 + c - 1] * source;
          I2(SigtVol, g, c, G) = I2(Sigt, g, zone, G) * Volume[c0 + c - 1];
        }
      for (int c = 0; c < MAXCORNER; ++c) nxez[c] = 0;

      for (int c = 1; c <= nCorner; ++c) {
        sumArea[c - 1] = 0.0;
        const int nCFaces = nCFacesArray[c0 + c - 1];
        for (int cface = 1; cface <= nCFaces; ++cface) {
          afp[cface - 1] = dot3_tol(omega, &I3(A_fp, 1, cface, c0 + c, 3, MAXCF), parTol);
          const int cfp = I2(cFP, cface, c0 + c, MAXCF);
          if (afp[cface - 1] > 0.0) {
            sumArea[c - 1] += afp[cface - 1];
            if (cfp > ncorner) {
              nxBdy += 1;
              bdy_exit[0][nxBdy - 1] = c;
              bdy_exit[1][nxBdy - 1] = cfp - ncorner;
            }
          } else if (afp[cface - 1] < 0.0) {
            for (int g = 1; g <= G; ++g) {
              I2(psifp, g, cface, G) = I2(Psi1, g, cfp, G);
              I2(src, g, c, G) = I2(src, g, c, G) - afp[cface - 1] * I2(Psi1, g, cfp, G);
            }
          }
        }
        for (int cface = 1; cface <= nCFaces; ++cface) {
          const double aez = dot3_tol(omega, &I3(A_ez, 1, cface, c0 + c, 3, MAXCF), parTol);
          const int cez = I2(cEZ, cface, c0 + c, MAXCF);
          if (cez > c) {
            if (aez > 0.0) {
              nxez[c - 1] += 1;
              ez_exit[nxez[c - 1] - 1][c - 1] = cez;
              coefpsi[nxez[c - 1] - 1][c - 1] = aez;
            } else if (aez < 0.0) {
              nxez[cez - 1] += 1;
              ez_exit[nxez[cez - 1] - 1][cez - 1] = c;
              coefpsi[nxez[cez - 1] - 1][cez - 1] = -aez;
            }
          }
          if (aez > 0.0) {
            sumArea[c - 1] += aez;
            double area_opp = 0.0;
            if (nCFaces == 3) {
              const int ifp = cface % nCFaces + 1;
              if (afp[ifp - 1] < 0.0) {
                for (int g = 1; g <= G; ++g) psi_opp[g - 1] = I2(psifp, g, ifp, G);
                area_opp = -afp[ifp - 1];
              }
            } else {
              int ifp = cface;
              area_opp = 0.0;
              for (int g = 0; g < G; ++g) psi_opp[g] = 0.0;
              for (int k = 1; k <= nCFaces - 2; ++k) {
                ifp = ifp % nCFaces + 1;
                if (afp[ifp - 1] < 0.0) {
                  area_opp = area_opp - afp[ifp - 1];
                  for (int g = 1; g <= G; ++g)
                    psi_opp[g - 1] = psi_opp[g - 1] - afp[ifp - 1] * I2(psifp, g, ifp, G);
                }
              }
              if (area_opp > 0.0) {
                const double area_inv = 1.0 / area_opp;
                for (int g = 0; g < G; ++g) psi_opp[g] = psi_opp[g] * area_inv;
              }
            }
            if (area_opp > 0.0) {
              const double aez2 = aez * aez;
              const double vol = Volume[c0 + c - 1];
              for (int g = 1; g <= G; ++g) {
                const double sig = I2(Sigt, g, zone, G);
                const double sigv = sig * vol;
                const double sigv2 = sigv * sigv;
                const double gnum = aez2 * (fouralpha * sigv2 + aez * (4.0 * sigv + 3.0 * aez));
                const double gden =
                    vol * (4.0 * sigv * sigv2 + aez * (6.0 * sigv2 + 2.0 * aez * (2.0 * sigv + aez)));
                const double sez = (vol * gnum * (sig * psi_opp[g - 1] - I2(Q, g, c, G)) +
                                    0.5 * aez * gden * (I2(Q, g, c, G) - I2(Q, g, cez, G))) /
                                   (gnum + gden * sig);
                I2(src, g, c, G) = I2(src, g, c, G) + sez;
                I2(src, g, cez, G) = I2(src, g, cez, G) - sez;
              }
            } else {
              for (int g = 1; g <= G; ++g) {
                const double sigInv = 1.0 / I2(Sigt, g, zone, G);
                const double sez = 0.5 * aez * sigInv * (I2(Q, g, c, G) - I2(Q, g, cez, G));
                I2(src, g, c, G) = I2(src, g, c, G) + sez;
                I2(src, g, cez, G) = I2(src, g, cez, G) - sez;
              }
            }
          }
        }
      }

      if (zone0 > 0) {
        for (int i = 1; i <= nCorner; ++i) {
          const int c = nextC[c0 + i - 1];
          for (int g = 1; g <= G; ++g) {
            I2(Psi1, g, c0 + c, G) = I2(src, g, c, G) / (sumArea[c - 1] + I2(SigtVol, g, c, G));
            I2(Phi, g, c0 + c, G) = I2(Phi, g, c0 + c, G) + quadwt * I2(Psi1, g, c0 + c, G);
          }
          for (int cface = 1; cface <= nxez[c - 1]; ++cface) {
            const int cez = ez_exit[cface - 1][c - 1];
            for (int g = 1; g <= G; ++g)
              I2(src, g, cez, G) = I2(src, g, cez, G) + coefpsi[cface - 1][c - 1] * I2(Psi1, g, c0 + c, G);
          }
        }
      } else {
        for (int c = 1; c <= nCorner; ++c)
          for (int cface = 1; cface <= nxez[c - 1]; ++cface) {
            const int cez = ez_exit[cface - 1][c - 1];
            for (int g = 1; g <= G; ++g)
              I2(src, g, cez, G) = I2(src, g, cez, G) + coefpsi[cface - 1][c - 1] * I2(Psi1, g, c0 + c, G);
          }
        for (int c = 1; c <= nCorner; ++c)
          for (int g = 1; g <= G; ++g) {
            I2(Psi1, g, c0 + c, G) = I2(src, g, c, G) / (sumArea[c - 1] + I2(SigtVol, g, c, G));
            I2(Phi, g, c0 + c, G) = I2(Phi, g, c0 + c, G) + quadwt * I2(Psi1, g, c0 + c, G);
          }
      }
      for (int ib = 1; ib <= nxBdy; ++ib) {
        const int c = bdy_exit[0][ib - 1];
        const int b = bdy_exit[1][ib - 1];
        for (int g = 1; g <= G; ++g) I2(PsiB, g, b, G) = I2(Psi1, g, c0 + c, G);
      }
    }
    ndoneZ += nz;
  }
  /* updateCycleList (control/constructDynMemory.F90:156-189) */
  for (int m = 1; m <= numCycles; ++m)
    for (int g = 1; g <= G; ++g) I2(cyclePsi, g, m, G) = I2(Psi1, g, cycleList[m - 1], G);
  free(psi_opp); free(SigtVol); free(src); free(Q); free(psifp);
}

void oracle_sweep_angle(int nzones, int ncorner, int nbelem, int Groups, const int *numCorner,
                        const int *cOffSet, const int *nCFacesArray, const int *cFP,
                        const int *cEZ, const double *A_fp, const double *A_ez,
                        const double *Volume, const double *STotal, const double *Sigt,
                        double tau, const double *Psi, const double *omega, double quadwt,
                        int nHyperPlanes, const int *zonesInPlane, const int *nextZ,
                        const int *nextC, double *Psi1, double *PsiB, double *Phi) {
  sweep_angle_tol(nzones, ncorner, nbelem, Groups, numCorner, cOffSet, nCFacesArray, cFP, cEZ, A_fp,
                  A_ez, Volume, STotal, Sigt, tau, Psi, omega, quadwt, nHyperPlanes, zonesInPlane,
                  nextZ, nextC, 0.0, 0, NULL, NULL, Psi1, PsiB, Phi);
}

void oracle_sweep_angle_cyc(int nzones, int ncorner, int nbelem, int Groups, const int *numCorner,
                            const int *cOffSet, const int *nCFacesArray, const int *cFP,
                            const int *cEZ, const double *A_fp, const double *A_ez,
                            const double *Volume, const double *STotal, const double *Sigt,
                            double tau, const double *Psi, const double *omega, double quadwt,
                            int nHyperPlanes, const int *zonesInPlane, const int *nextZ,
                            const int *nextC, int numCycles, const int *cycleList,
                            double *cyclePsi, double *Psi1, double *PsiB, double *Phi) {
  sweep_angle_tol(nzones, ncorner, nbelem, Groups, numCorner, cOffSet, nCFacesArray, cFP, cEZ, A_fp,
                  A_ez, Volume, STotal, Sigt, tau, Psi, omega, quadwt, nHyperPlanes, zonesInPlane,
                  nextZ, nextC, 0.0, numCycles, cycleList, cyclePsi, Psi1, PsiB, Phi);
}

/* ------------------------------------------------------------------------ */
/* snac/SetSweep.F90:100-170 (single domain, vacuum: one flux iteration) */
void oracle_sweep_all(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                      const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                      const double *A_fp, const double *A_ez, const double *Volume,
                      const double *STotal, const double *Sigt, const double *omega,
                      const double *weight, const int *nHyperPlanes, const int *zonesInPlane,
                      const int *nextZ, const int *nextC, int nthreads, double *PsiB,
                      double *Phi) {
  oracle_sweep_all_tol(nzones, ncorner, nbelem, G, A, numCorner, cOffSet, nCFaces, cFP, cEZ, A_fp,
                       A_ez, Volume, STotal, Sigt, omega, weight, nHyperPlanes, zonesInPlane, nextZ,
                       nextC, nthreads, 0.0, PsiB, Phi);
}

void oracle_sweep_all_tol(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                          const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                          const double *A_fp, const double *A_ez, const double *Volume,
                          const double *STotal, const double *Sigt, const double *omega,
                          const double *weight, const int *nHyperPlanes, const int *zonesInPlane,
                          const int *nextZ, const int *nextC, int nthreads, double parTol,
                          double *PsiB, double *Phi) {
  oracle_sweep_all_cyc(nzones, ncorner, nbelem, G, A, numCorner, cOffSet, nCFaces, cFP, cEZ, A_fp,
                       A_ez, Volume, STotal, Sigt, omega, weight, nHyperPlanes, zonesInPlane, nextZ,
                       nextC, nthreads, parTol, NULL, NULL, NULL, PsiB, Phi);
}

void oracle_sweep_all_cyc(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                          const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                          const double *A_fp, const double *A_ez, const double *Volume,
                          const double *STotal, const double *Sigt, const double *omega,
                          const double *weight, const int *nHyperPlanes, const int *zonesInPlane,
                          const int *nextZ, const int *nextC, int nthreads, double parTol,
                          const int *numCycles, const int *cycleList, double *cyclePsi,
                          double *PsiB, double *Phi) {
  const size_t nGC = (size_t)G * ncorner, nPsi1 = (size_t)G * (ncorner + nbelem);
  int *planeStart = malloc(sizeof(int) * (A + 1));
  int *cycleOffSet = malloc(sizeof(int) * (A + 1)); /* rt/rtorder.F90:62-69 */
  planeStart[0] = 0;
  cycleOffSet[0] = 0;
  for (int a = 0; a < A; ++a) {
    planeStart[a + 1] = planeStart[a] + nHyperPlanes[a];
    cycleOffSet[a + 1] = cycleOffSet[a] + (numCycles ? numCycles[a] : 0);
  }
  if (nthreads < 1) nthreads = 1;
  /* One phase-space set per thread, each with its own Phi (Set%Phi), summed
   * afterwards in set order like getPhiTotal (control/getPhiTotal_OMPOL.F90:134-162). */
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
    double *phiT = phiSets + (size_t)t * nGC;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int a = 0; a < A; ++a) {
      const int nc = numCycles ? numCycles[a] : 0;
      sweep_angle_tol(nzones, ncorner, nbelem, G, numCorner, cOffSet, nCFaces, cFP, cEZ, A_fp,
                      A_ez, Volume, STotal, Sigt, 0.0, NULL, &omega[3 * (size_t)a], weight[a],
                      nHyperPlanes[a], &zonesInPlane[planeStart[a]], &nextZ[(size_t)a * nzones],
                      &nextC[(size_t)a * ncorner], parTol, nc,
                      nc ? &cycleList[cycleOffSet[a]] : NULL,
                      nc ? &cyclePsi[(size_t)G * cycleOffSet[a]] : NULL, Psi1,
                      &PsiB[(size_t)a * G * nbelem], phiT);
    }
    free(Psi1);
  }
  for (size_t i = 0; i < nGC; ++i) {
    double s = 0.0;
    for (int t = 0; t < nthreads; ++t) s += phiSets[(size_t)t * nGC + i];
    Phi[i] = s;
  }
  free(phiSets);
  free(planeStart);
  free(cycleOffSet);
}

/* ------------------------------------------------------------------------ */
int oracle_source_iteration(int nzones, int ncorner, int nbelem, int G, int A,
                            const int *numCorner, const int *cOffSet, const int *nCFaces,
                            const int *cFP, const int *cEZ, const double *A_fp,
                            const double *A_ez, const double *Volume, const double *VolumeZone,
                            const double *Sigt, const double *sigs, const double *q,
                            const double *omega, const double *weight, const int *nHyperPlanes,
                            const int *zonesInPlane, const int *nextZ, const int *nextC,
                            int nthreads, double tol, int maxIters, double *Phi,
                            double *residual, double *hist) {
  return oracle_source_iteration_cyc(nzones, ncorner, nbelem, G, A, numCorner, cOffSet, nCFaces,
                                     cFP, cEZ, A_fp, A_ez, Volume, VolumeZone, Sigt, sigs, q,
                                     omega, weight, nHyperPlanes, zonesInPlane, nextZ, nextC,
                                     NULL, NULL, NULL, nthreads, tol, maxIters, Phi, residual,
                                     hist);
}

int oracle_source_iteration_cyc(int nzones, int ncorner, int nbelem, int G, int A,
                                const int *numCorner, const int *cOffSet, const int *nCFaces,
                                const int *cFP, const int *cEZ, const double *A_fp,
                                const double *A_ez, const double *Volume, const double *VolumeZone,
                                const double *Sigt, const double *sigs, const double *q,
                                const double *omega, const double *weight, const int *nHyperPlanes,
                                const int *zonesInPlane, const int *nextZ, const int *nextC,
                                const int *numCycles, const int *cycleList, double *cyclePsi,
                                int nthreads, double tol, int maxIters, double *Phi,
                                double *residual, double *hist) {
  const size_t nGC = (size_t)G * ncorner, nGZ = (size_t)G * nzones;
  const double wtiso = 1.0 / (4.0 * M_PI);
  double *STotal = malloc(sizeof(double) * nGC);
  double *PsiB = calloc((size_t)G * nbelem * A, sizeof(double));
  double *old = malloc(sizeof(double) * nGZ), *cur = malloc(sizeof(double) * nGZ);
  /* zone averages of the starting guess */
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
    memset(PsiB, 0, sizeof(double) * (size_t)G * nbelem * A); /* vacuum */
    oracle_sweep_all_cyc(nzones, ncorner, nbelem, G, A, numCorner, cOffSet, nCFaces, cFP, cEZ,
                         A_fp, A_ez, Volume, STotal, Sigt, omega, weight, nHyperPlanes,
                         zonesInPlane, nextZ, nextC, nthreads, 0.0, numCycles, cycleList, cyclePsi,
                         PsiB, Phi);
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
