// Per-angle sweep schedule: zone "need" counts and exit faces
// (snac/snneed.F90:95-191), seeds (snac/findseeds.F90:58-66), corner order
// inside each zone (snac/getDownStreamData.F90:88-149) and the hyperplane
// layering (snac/snnext.F90:103-195), called per angle as rt/rtorder.F90:54-60
// does.
//
// Meshes with dependency cycles (warped faces, cyclic zones or zone cycles)
// are rejected with an error: cycle breaking (snac/cyclebreaker.F90,
// snac/sccsearch.F90) and lagged cycle fluxes are not built here.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "host.h"

namespace umt {

namespace {

inline double dot3(const double* o, const double* a) {
  return o[0] * a[0] + o[1] * a[1] + o[2] * a[2];
}

// Omega.A, or exactly 0 when the corner face is parallel to Omega within
// tol * |A|_1 (directions lying in a mesh face leave rounding noise of
// either sign; with tol = 0 this is the plain dot product).
inline double face_dot(const double* o, const double* a, double tol) {
  const double d = dot3(o, a);
  if (tol > 0.0 && std::fabs(d) <= tol * (std::fabs(a[0]) + std::fabs(a[1]) + std::fabs(a[2])))
    return 0.0;
  return d;
}

// One angle. Returns 0 or an error code with a message.
int schedule_angle(const Mesh& m, const Geometry& g, const double* omega, int angle,
                   Schedule& s, std::vector<int>& planes, std::string& err, double tol) {
  const int Z = m.nzones;
  std::vector<int> needZ(Z, 0);
  std::vector<char> exitFace(kMaxFaces * (size_t)Z, 0);
  std::vector<char> onCycle(Z, 0);
  int meshCycles = 0;

  // snneed: face-summed Omega.A_fp decides the zone dependency across a face;
  // corner faces of one face with mixed signs put a zone on the cycle list.
  for (int z = 0; z < Z; ++z) {
    const int c0 = m.cOffSet[z], nCorner = m.numCorner[z], nFaces = m.zoneFaces[z];
    double afpmFace[kMaxFaces] = {0};
    int nInc[kMaxFaces] = {0}, nExit[kMaxFaces] = {0};
    for (int c = 0; c < nCorner; ++c) {
      const size_t cc = c0 + c;
      for (int cf = 0; cf < m.nCFaces[cc]; ++cf) {
        const int face = m.CToFace[kMaxCF * cc + cf] - 1;
        const int zoneOpp = m.zoneOpp[kMaxFaces * (size_t)z + face];
        if (zoneOpp > z + 1) {
          const double afpm = face_dot(omega, &g.A_fp[3 * (kMaxCF * cc + cf)], tol);
          afpmFace[face] += afpm;
          if (afpm < 0.0) nInc[face] += 1;
          else if (afpm > 0.0) nExit[face] += 1;
        }
      }
    }
    for (int f = 0; f < nFaces; ++f) {
      const int zoneOpp = m.zoneOpp[kMaxFaces * (size_t)z + f];
      if (zoneOpp > z + 1) {
        const int faceOpp = m.faceOpp[kMaxFaces * (size_t)z + f];
        if (afpmFace[f] < 0.0) {
          needZ[z] += 1;
          exitFace[kMaxFaces * (size_t)(zoneOpp - 1) + faceOpp - 1] = 1;
          if (nExit[f] > 0 && !onCycle[z]) { onCycle[z] = 1; meshCycles += nCorner; }
        } else if (afpmFace[f] > 0.0) {
          needZ[zoneOpp - 1] += 1;
          exitFace[kMaxFaces * (size_t)z + f] = 1;
          if (nInc[f] > 0 && !onCycle[zoneOpp - 1]) {
            onCycle[zoneOpp - 1] = 1;
            meshCycles += m.numCorner[zoneOpp - 1];
          }
        }
      }
    }
  }
  if (meshCycles > 0) {
    char buf[160];
    snprintf(buf, sizeof buf, "angle %d: %d corners on warped-face cycles (unsupported)", angle + 1,
             meshCycles);
    err = buf;
    return 2;
  }

  // findseeds
  std::vector<int> listZone;
  listZone.reserve(Z);
  for (int z = 0; z < Z; ++z)
    if (needZ[z] == 0) listZone.push_back(z);
  if (listZone.empty()) {
    err = "no seed zones (mesh cycle at the boundary, unsupported)";
    return 3;
  }

  // getDownStreamData: lowest-index ready corner first (minloc of need).
  for (int z = 0; z < Z; ++z) {
    const int c0 = m.cOffSet[z], nCorner = m.numCorner[z];
    int need[kMaxCorner] = {0}, nDSC[kMaxCorner] = {0};
    int ds[kMaxCorner][kMaxCF];
    for (int c = 0; c < nCorner; ++c) {
      const size_t cc = c0 + c;
      for (int cf = 0; cf < m.nCFaces[cc]; ++cf) {
        const int cez = m.cEZ[kMaxCF * cc + cf];  // 1-based local
        if (cez > c + 1) {
          const double aez = face_dot(omega, &g.A_ez[3 * (kMaxCF * cc + cf)], tol);
          if (aez < 0.0) {
            need[c] += 1;
            ds[cez - 1][nDSC[cez - 1]++] = c;
          } else if (aez > 0.0) {
            need[cez - 1] += 1;
            ds[c][nDSC[c]++] = cez - 1;
          }
        }
      }
    }
    bool bad = false;
    for (int i = 0; i < nCorner; ++i) {
      int c = 0;
      for (int k = 1; k < nCorner; ++k)
        if (need[k] < need[c]) c = k;
      if (need[c] != 0) bad = true;
      s.nextC[(size_t)angle * m.ncorner + c0 + i] = c + 1;
      for (int k = 0; k < nDSC[c]; ++k) need[ds[c][k]] -= 1;
      need[c] = 99;
    }
    if (bad) {
      char buf[160];
      snprintf(buf, sizeof buf, "angle %d: zone %d has a corner dependency cycle (unsupported)",
               angle + 1, z + 1);
      err = buf;
      return 4;
    }
  }

  // snnext: breadth-first layering into hyperplanes.
  std::vector<char> done(Z, 0);
  int ndone = 0, lastZone = 0, newZones = (int)listZone.size();
  planes.clear();
  int* nextZ = &s.nextZ[(size_t)angle * Z];
  while (true) {
    planes.push_back(newZones);
    int added = 0;
    for (int zid = 0; zid < newZones; ++zid) {
      const int z = listZone[lastZone + zid];
      done[z] = 1;
      for (int f = 0; f < m.zoneFaces[z]; ++f) {
        if (!exitFace[kMaxFaces * (size_t)z + f]) continue;
        const int zx = m.zoneOpp[kMaxFaces * (size_t)z + f];
        if (zx > 0 && !done[zx - 1]) {
          if (--needZ[zx - 1] == 0) {
            listZone.push_back(zx - 1);
            ++added;
          } else if (needZ[zx - 1] < 0) {
            err = "needZ < 0 while layering";
            return 5;
          }
        }
      }
      nextZ[ndone++] = z + 1;
    }
    lastZone += newZones;
    if (lastZone == Z) break;
    if (added == 0) {
      char buf[160];
      snprintf(buf, sizeof buf, "angle %d: zone dependency cycle after %d zones (unsupported)",
               angle + 1, lastZone);
      err = buf;
      return 6;
    }
    newZones = added;
  }
  return 0;
}

}  // namespace

int build_schedule(const Mesh& m, const Geometry& g, const Quadrature& q, Schedule& s,
                   std::string& err, double parallelTol) {
  const int A = q.nangles;
  s = Schedule();
  s.nangles = A;
  s.nzones = m.nzones;
  s.ncorner = m.ncorner;
  s.nHyperPlanes.assign(A, 0);
  s.numCycles.assign(A, 0);
  s.nextZ.assign((size_t)A * m.nzones, 0);
  s.nextC.assign((size_t)A * m.ncorner, 0);
  std::vector<std::vector<int>> planes(A);
  int rc = 0;
  std::string firstErr;
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < A; ++a) {
    std::string e;
    int r = schedule_angle(m, g, &q.omega[3 * (size_t)a], a, s, planes[a], e, parallelTol);
    if (r) {
#pragma omp critical
      if (!rc) { rc = r; firstErr = e; }
    }
  }
  if (rc) { err = firstErr; return rc; }
  s.planeOffset.assign(A + 1, 0);
  for (int a = 0; a < A; ++a) {
    s.nHyperPlanes[a] = (int)planes[a].size();
    s.planeOffset[a + 1] = s.planeOffset[a] + s.nHyperPlanes[a];
    s.maxPlanes = std::max(s.maxPlanes, s.nHyperPlanes[a]);
  }
  s.zonesInPlane.resize(s.planeOffset[A]);
  for (int a = 0; a < A; ++a)
    std::copy(planes[a].begin(), planes[a].end(), s.zonesInPlane.begin() + s.planeOffset[a]);
  return 0;
}

}  // namespace umt
