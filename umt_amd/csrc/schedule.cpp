// Per-angle sweep schedule: zone "need" counts and exit faces
// (snac/snneed.F90:95-191), seeds (snac/findseeds.F90:58-66), corner order
// inside each zone (snac/getDownStreamData.F90:88-149) and the hyperplane
// layering (snac/snnext.F90:103-195), called per angle as rt/rtorder.F90:54-60
// does.
//
// Meshes with dependency cycles are scheduled as the reference schedules
// them: warped faces with mixed-sign corner faces, zones whose corners
// depend on each other in a cycle (nextZ < 0) and cycles among zones
// (snac/cyclebreaker.F90, snac/sccsearch.F90) put the corners whose psi is
// read before it is computed on the angle's cycle list; the sweep reads the
// previous sweep's values there.
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

// Dependency-graph state of one angle while it is layered.
struct AngleGraph {
  const Mesh& m;
  std::vector<int> needZ;       // unmet upstream zones
  std::vector<char> exitFace;   // (kMaxFaces, Z): flow leaves the zone through the face
  std::vector<char> onCycle;    // zone's corners are already on the cycle list
  std::vector<int>& cycles;     // 1-based corners whose previous-sweep psi is read
  explicit AngleGraph(const Mesh& mesh, std::vector<int>& cyc)
      : m(mesh), needZ(mesh.nzones, 0), exitFace(kMaxFaces * (size_t)mesh.nzones, 0),
        onCycle(mesh.nzones, 0), cycles(cyc) {}
  char& exit(int z, int f) { return exitFace[kMaxFaces * (size_t)z + f]; }
  int opp(int z, int f) const { return m.zoneOpp[kMaxFaces * (size_t)z + f]; }      // 1-based, < 0 boundary
  int oppFace(int z, int f) const { return m.faceOpp[kMaxFaces * (size_t)z + f]; }  // 1-based
  void add_zone(int z) {
    for (int c = 1; c <= m.numCorner[z]; ++c) cycles.push_back(m.cOffSet[z] + c);
  }
};

// Breaks the dependency cycles among the zones not yet layered
// (snac/cyclebreaker.F90:75-150): a Tarjan strongly-connected-component
// search (snac/sccsearch.F90) from every remaining zone in zone order, faces
// in face order. When a component of more than one zone closes, every
// exit link into its root from a zone still on the stack is removed, the
// upstream zone's corners go on the cycle list (its psi is read lagged), and
// the root is released when its need count reaches zero. The search is an
// explicit-stack version of the recursive one, visiting the same zones and
// faces in the same order. Released roots are appended to listZone; returns
// how many, or -1 with a message.
int break_cycles(AngleGraph& gr, int ndone, std::vector<int>& listZone, std::string& err) {
  const Mesh& m = gr.m;
  const int Z = m.nzones;
  const int ngraph = Z - ndone;
  std::vector<char> fresh(Z, 1), onstack(Z, 0);
  std::vector<int> dfnum(Z, 0), lowlink(Z, 0), stack, temp, released;
  std::vector<int> listZ;
  for (int z = 0; z < Z; ++z) {
    if (gr.needZ[z] == 0) fresh[z] = 0;
    else listZ.push_back(z);
  }
  if ((int)listZ.size() != ngraph) {
    err = "miscount of remaining zones while breaking cycles";
    return -1;
  }
  int ncount = 0;
  struct Frame {
    int zone, face, child;
  };
  std::vector<Frame> frames;
  auto enter = [&](int z) {
    dfnum[z] = lowlink[z] = ++ncount;
    fresh[z] = 0;
    stack.push_back(z);
    onstack[z] = 1;
    frames.push_back({z, 0, -1});
  };
  for (int start : listZ) {
    if (!fresh[start]) continue;
    enter(start);
    while (!frames.empty()) {
      Frame& fr = frames.back();
      const int z = fr.zone;
      if (fr.child >= 0) {  // back from the search of a downstream zone
        if (lowlink[fr.child] < lowlink[z]) lowlink[z] = lowlink[fr.child];
        fr.child = -1;
      }
      bool descended = false;
      while (fr.face < m.zoneFaces[z]) {
        const int f = fr.face++;
        if (!gr.exit(z, f)) continue;
        const int z2 = gr.opp(z, f) - 1;
        if (z2 < 0) continue;
        if (fresh[z2]) {
          fr.child = z2;
          enter(z2);  // invalidates fr
          descended = true;
          break;
        }
        if (dfnum[z2] < dfnum[z] && onstack[z2] && lowlink[z2] < lowlink[z]) lowlink[z] = lowlink[z2];
      }
      if (descended) continue;
      // All faces seen: close a component rooted here.
      if (lowlink[z] == dfnum[z]) {
        int z2 = stack.back();
        stack.pop_back();
        onstack[z2] = 0;
        if (z2 != z) {
          temp.clear();
          while (z2 != z) {
            temp.push_back(z2);
            z2 = stack.back();
            stack.pop_back();
          }
          temp.push_back(z2);
          onstack[temp[0]] = 1;  // every member is on the stack while links are broken
          const int root = temp.back();
          for (int f = 0; f < m.zoneFaces[root]; ++f) {
            const int zb = gr.opp(root, f) - 1;
            if (zb < 0) continue;
            const int fb = gr.oppFace(root, f) - 1;
            if (onstack[zb] && gr.exit(zb, fb)) {
              if (!gr.onCycle[zb]) {
                gr.add_zone(zb);
                gr.onCycle[zb] = 1;
              }
              gr.needZ[root] -= 1;
              gr.exit(zb, fb) = 0;
              if (gr.needZ[root] == 0) released.push_back(root);
            }
          }
          for (int t : temp) onstack[t] = 0;
        }
      }
      frames.pop_back();
    }
  }
  if (released.empty()) {
    err = "cycle detection failed: no dependency broken";
    return -1;
  }
  int added = 0;
  for (int z : released) {
    if (gr.needZ[z] == 0) {
      listZone.push_back(z);
      ++added;
    } else if (gr.needZ[z] < 0) {
      err = "negative need count after breaking cycles";
      return -1;
    }
  }
  if (added == 0) {
    err = "cycles found but not broken";
    return -1;
  }
  return added;
}

// One angle. Returns 0 or an error code with a message.
int schedule_angle(const Mesh& m, const Geometry& g, const double* omega, int angle,
                   Schedule& s, std::vector<int>& planes, std::vector<int>& cycles,
                   std::string& err, double tol) {
  const int Z = m.nzones;
  cycles.clear();
  AngleGraph gr(m, cycles);

  // snneed: face-summed Omega.A_fp decides the zone dependency across a face;
  // corner faces of one face with mixed signs put the downstream zone on the
  // cycle list (some of its corners feed the upstream zone).
  for (int z = 0; z < Z; ++z) {
    const int c0 = m.cOffSet[z], nCorner = m.numCorner[z], nFaces = m.zoneFaces[z];
    double afpmFace[kMaxFaces] = {0};
    int nInc[kMaxFaces] = {0}, nExit[kMaxFaces] = {0};
    for (int c = 0; c < nCorner; ++c) {
      const size_t cc = c0 + c;
      for (int cf = 0; cf < m.nCFaces[cc]; ++cf) {
        const int face = m.CToFace[kMaxCF * cc + cf] - 1;
        const int zoneOpp = gr.opp(z, face);
        if (zoneOpp > z + 1) {
          const double afpm = face_dot(omega, &g.A_fp[3 * (kMaxCF * cc + cf)], tol);
          afpmFace[face] += afpm;
          if (afpm < 0.0) nInc[face] += 1;
          else if (afpm > 0.0) nExit[face] += 1;
        }
      }
    }
    for (int f = 0; f < nFaces; ++f) {
      const int zoneOpp = gr.opp(z, f);
      if (zoneOpp > z + 1) {
        const int faceOpp = gr.oppFace(z, f);
        if (afpmFace[f] < 0.0) {
          gr.needZ[z] += 1;
          gr.exit(zoneOpp - 1, faceOpp - 1) = 1;
          if (nExit[f] > 0 && !gr.onCycle[z]) {
            gr.add_zone(z);
            gr.onCycle[z] = 1;
          }
        } else if (afpmFace[f] > 0.0) {
          gr.needZ[zoneOpp - 1] += 1;
          gr.exit(z, f) = 1;
          if (nInc[f] > 0 && !gr.onCycle[zoneOpp - 1]) {
            gr.add_zone(zoneOpp - 1);
            gr.onCycle[zoneOpp - 1] = 1;
          }
        }
      }
    }
  }

  // findseeds: zones with no upstream zone; with none, the boundary zone of
  // least need is forced to be a seed and its upstream neighbours are read
  // lagged (snac/findseeds.F90:58-114).
  std::vector<int> listZone;
  listZone.reserve(Z);
  for (int z = 0; z < Z; ++z)
    if (gr.needZ[z] == 0) listZone.push_back(z);
  if (listZone.empty()) {
    int seed = -1, minNeed = Z;
    for (int z = 0; z < Z; ++z) {
      bool boundary = false;
      for (int f = 0; f < m.zoneFaces[z]; ++f) boundary |= gr.opp(z, f) < 0;
      if (boundary && gr.needZ[z] < minNeed) {
        seed = z;
        minNeed = gr.needZ[z];
      }
    }
    if (seed < 0) {
      err = "no seed zone";
      return 3;
    }
    listZone.push_back(seed);
    gr.needZ[seed] = 0;
    for (int f = 0; f < m.zoneFaces[seed]; ++f) {
      if (gr.exit(seed, f)) continue;
      const int zo = gr.opp(seed, f);
      if (zo > 0) {
        gr.add_zone(zo - 1);
        gr.exit(zo - 1, gr.oppFace(seed, f) - 1) = 0;
        gr.onCycle[zo - 1] = 1;
      }
    }
  }

  // getDownStreamData: lowest-index ready corner first (minloc of need). A
  // zone whose corners depend on each other in a cycle keeps the natural
  // corner order, is solved from lagged values of its own corners (nextZ < 0)
  // and has its corners put on the cycle list (snac/fixZone.F90).
  std::vector<char> bad(Z, 0);
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
    int* nextC = &s.nextC[(size_t)angle * m.ncorner + c0];
    for (int i = 0; i < nCorner; ++i) {
      int c = 0;
      for (int k = 1; k < nCorner; ++k)
        if (need[k] < need[c]) c = k;
      if (need[c] != 0) bad[z] = 1;
      nextC[i] = c + 1;
      for (int k = 0; k < nDSC[c]; ++k) need[ds[c][k]] -= 1;
      need[c] = 99;
    }
    if (bad[z]) {
      for (int i = 0; i < nCorner; ++i) nextC[i] = i + 1;
      gr.add_zone(z);
    }
  }

  // snnext: breadth-first layering into hyperplanes; when no zone becomes
  // ready, cycles among the remaining zones are broken.
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
        if (!gr.exit(z, f)) continue;
        const int zx = gr.opp(z, f);
        if (zx > 0 && !done[zx - 1]) {
          if (--gr.needZ[zx - 1] == 0) {
            listZone.push_back(zx - 1);
            ++added;
          } else if (gr.needZ[zx - 1] < 0) {
            err = "needZ < 0 while layering";
            return 5;
          }
        }
      }
      nextZ[ndone++] = bad[z] ? -(z + 1) : z + 1;
    }
    lastZone += newZones;
    if (lastZone == Z) break;
    if (added == 0) {
      added = break_cycles(gr, ndone, listZone, err);
      if (added < 0) {
        char buf[200];
        snprintf(buf, sizeof buf, "angle %d: %s", angle + 1, err.c_str());
        err = buf;
        return 6;
      }
    }
    newZones = added;
  }
  if ((long)cycles.size() > (long)m.ncorner) {
    char buf[160];
    snprintf(buf, sizeof buf, "angle %d: %zu cycle corners exceed the %d corners", angle + 1,
             cycles.size(), m.ncorner);
    err = buf;
    return 7;
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
  std::vector<std::vector<int>> planes(A), cycles(A);
  int rc = 0;
  std::string firstErr;
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < A; ++a) {
    std::string e;
    int r = schedule_angle(m, g, &q.omega[3 * (size_t)a], a, s, planes[a], cycles[a], e,
                           parallelTol);
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
  // Per-angle cycle lists concatenated in angle order (rt/rtorder.F90:62-90).
  for (int a = 0; a < A; ++a) {
    s.numCycles[a] = (int)cycles[a].size();
    s.cycleList.insert(s.cycleList.end(), cycles[a].begin(), cycles[a].end());
  }
  return 0;
}

}  // namespace umt
