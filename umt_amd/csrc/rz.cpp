// 2D cylindrical (RZ) problem preparation on the host: quad-mesh corner
// topology (aux/setTetonZone.F90:122-142, aux/setOppositeFace.F90:32-53),
// corner geometry (rt/geometryUCBrz.F90), the RZ product quadrature
// (rt/quadrz.F90:184-268, rt/rtquad.F90:80-123) with its angular
// coefficients (rt/AngleCoef2D.F90, mods/AngleSet_mod.F90:312-347), the 2D
// sweep schedule (snac/snneed.F90:62-93, findseeds, getDownStreamData,
// snnext) and the device sweep levels.
#include "rz.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <stdexcept>

#include "qrs45.h"

namespace umt {

void build_rz_mesh(int npoints, const std::vector<double>& points,
                   const std::vector<std::array<int, 4>>& quadsIn, RzMesh& m) {
  const int Z = (int)quadsIn.size();
  if (Z < 1 || npoints < 4) throw std::invalid_argument("empty quad mesh");
  m = RzMesh();
  m.nzones = Z;
  m.ncorner = 4 * Z;
  m.npoints = npoints;
  m.points = points;
  m.quads = quadsIn;
  auto R = [&](int n) { return points[2 * (size_t)n]; };
  auto Zc = [&](int n) { return points[2 * (size_t)n + 1]; };
  for (auto& q : m.quads) {
    for (int k = 0; k < 4; ++k)
      if (q[k] < 0 || q[k] >= npoints) throw std::invalid_argument("node id out of range");
    double a2 = 0.0;
    for (int k = 0; k < 4; ++k) a2 += R(q[k]) * Zc(q[(k + 1) % 4]) - R(q[(k + 1) % 4]) * Zc(q[k]);
    if (!(a2 != 0.0) || !std::isfinite(a2)) throw std::invalid_argument("degenerate quad zone");
    if (a2 > 0.0) std::swap(q[1], q[3]);  // counterclockwise in (r, z): reverse
  }
  const int C = m.ncorner;
  m.numCorner.assign(Z, 4);
  m.zoneFaces.assign(Z, 4);
  m.cOffSet.resize(Z);
  for (int z = 0; z < Z; ++z) m.cOffSet[z] = 4 * z;
  // Faces (edges) k = (node k, node k+1): opposite zones and faces.
  std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> edges;
  for (int z = 0; z < Z; ++z)
    for (int f = 0; f < 4; ++f) {
      const int a = m.quads[z][f], b = m.quads[z][(f + 1) % 4];
      edges[{std::min(a, b), std::max(a, b)}].push_back({z, f});
    }
  m.zoneOpp.assign(4 * (size_t)Z, -1);
  for (auto& e : edges) {
    if (e.second.size() > 2) throw std::invalid_argument("edge shared by more than two zones");
    if (e.second.size() == 2) {
      const auto [z1, f1] = e.second[0];
      const auto [z2, f2] = e.second[1];
      if (z1 == z2) throw std::invalid_argument("zone shares an edge with itself");
      m.zoneOpp[4 * (size_t)z1 + f1] = z2 + 1;
      m.zoneOpp[4 * (size_t)z2 + f2] = z1 + 1;
    }
  }
  m.faceOpp.assign(4 * (size_t)Z, 0);
  for (int z = 0; z < Z; ++z)
    for (int f = 0; f < 4; ++f) {
      const int zo = m.zoneOpp[4 * (size_t)z + f];
      if (zo < 0) {
        m.faceOpp[4 * (size_t)z + f] = -1;
        continue;
      }
      for (int f2 = 0; f2 < 4; ++f2)
        if (m.zoneOpp[4 * (size_t)(zo - 1) + f2] == z + 1) m.faceOpp[4 * (size_t)z + f] = f2 + 1;
    }
  // Corner faces: corner k's face 2 and corner k+1's face 1 lie on edge k.
  m.cFP.assign(2 * (size_t)C, 0);
  m.cEZ.assign(2 * (size_t)C, 0);
  m.CToFace.assign(2 * (size_t)C, 0);
  m.px.assign(2 * (size_t)C, 0.0);
  auto cornerAt = [&](int z, int node) {
    for (int k = 0; k < 4; ++k)
      if (m.quads[z][k] == node) return 4 * z + k + 1;
    throw std::invalid_argument("neighbouring zones do not share the edge's nodes");
  };
  int nb = 0;
  for (int z = 0; z < Z; ++z) {
    for (int k = 0; k < 4; ++k) {
      m.px[2 * (size_t)(4 * z + k)] = R(m.quads[z][k]);
      m.px[2 * (size_t)(4 * z + k) + 1] = Zc(m.quads[z][k]);
    }
    for (int f = 0; f < 4; ++f) {
      const int l1 = f, l2 = (f + 1) % 4;
      const int a = m.quads[z][f], b = m.quads[z][(f + 1) % 4];
      int opp1, opp2;
      const int zo = m.zoneOpp[4 * (size_t)z + f];
      if (zo < 0) {
        opp1 = C + nb + 1;
        opp2 = C + nb + 2;
        nb += 2;
      } else {
        opp1 = cornerAt(zo - 1, a);
        opp2 = cornerAt(zo - 1, b);
      }
      const size_t c1 = 4 * (size_t)z + l1, c2 = 4 * (size_t)z + l2;
      m.cFP[2 * c1 + 1] = opp1;
      m.cFP[2 * c2 + 0] = opp2;
      m.cEZ[2 * c1 + 0] = l2 + 1;
      m.cEZ[2 * c2 + 1] = l1 + 1;
      m.CToFace[2 * c1 + 1] = f + 1;
      m.CToFace[2 * c2 + 0] = f + 1;
    }
  }
  m.nbelem = nb;
}

void build_rz_structured(int nr, int nz, double r0, double r1, double z0, double z1,
                         double jitter, uint64_t seed, RzMesh& m) {
  if (nr < 1 || nz < 1 || !(r1 > r0) || !(z1 > z0) || r0 < 0.0 || !(jitter >= 0.0) ||
      jitter >= 0.5)
    throw std::invalid_argument("structured RZ mesh: bad extents or jitter (0 <= jitter < 0.5)");
  const int np = (nr + 1) * (nz + 1);
  std::vector<double> pts(2 * (size_t)np);
  const double dr = (r1 - r0) / nr, dz = (z1 - z0) / nz;
  for (int j = 0; j <= nz; ++j)
    for (int i = 0; i <= nr; ++i) {
      const int n = j * (nr + 1) + i;
      double r = i == nr ? r1 : r0 + i * dr, z = j == nz ? z1 : z0 + j * dz;
      if (jitter > 0.0 && i > 0 && i < nr && j > 0 && j < nz) {
        r += (2.0 * draw_uniform(seed, (uint64_t)n, 0, 11) - 1.0) * jitter * dr;
        z += (2.0 * draw_uniform(seed, (uint64_t)n, 0, 12) - 1.0) * jitter * dz;
      }
      pts[2 * (size_t)n] = r;
      pts[2 * (size_t)n + 1] = z;
    }
  std::vector<std::array<int, 4>> quads;
  quads.reserve((size_t)nr * nz);
  auto node = [&](int i, int j) { return j * (nr + 1) + i; };
  for (int j = 0; j < nz; ++j)
    for (int i = 0; i < nr; ++i)
      quads.push_back({node(i, j), node(i, j + 1), node(i + 1, j + 1), node(i + 1, j)});
  build_rz_mesh(np, pts, quads, m);
}

// Corner geometry as rt/geometryUCBrz.F90:45-157 computes it, face by face
// in the same order (a face shared by two corners is computed once and
// negated for the other), so the arrays match the reference's bit for bit.
void compute_rz_geometry(const RzMesh& m, RzGeometry& g) {
  const int Z = m.nzones, C = m.ncorner;
  const double half = 0.5, third = 1.0 / 3.0;
  g.A_fp.assign(4 * (size_t)C, 0.0);
  g.A_ez.assign(4 * (size_t)C, 0.0);
  g.RadiusFP.assign(2 * (size_t)C, 0.0);
  g.RadiusEZ.assign(2 * (size_t)C, 0.0);
  g.Area.assign(C, 0.0);
  g.Volume.assign(C, 0.0);
  g.VolumeZone.assign(Z, 0.0);
  auto AFP = [&](int d, int f, int c) -> double& { return g.A_fp[d + 2 * (f + 2 * (size_t)c)]; };
  auto AEZ = [&](int d, int f, int c) -> double& { return g.A_ez[d + 2 * (f + 2 * (size_t)c)]; };
  for (int z = 0; z < Z; ++z) {
    const int c0 = m.cOffSet[z], nCorner = m.numCorner[z];
    double zc0 = 0.0, zc1 = 0.0;
    for (int c = 0; c < nCorner; ++c) {
      zc0 += m.px[2 * (size_t)(c0 + c)];
      zc1 += m.px[2 * (size_t)(c0 + c) + 1];
    }
    zc0 /= (double)nCorner;
    zc1 /= (double)nCorner;
    for (int c = 1; c <= nCorner; ++c) {  // local corners 1-based, as the reference loops
      const int cc = c0 + c - 1;
      const int c1 = m.cEZ[2 * (size_t)cc], c2 = m.cEZ[2 * (size_t)cc + 1];
      const int cfp1 = m.cFP[2 * (size_t)cc], cfp2 = m.cFP[2 * (size_t)cc + 1];
      const double r_zone = zc0, z_zone = zc1;
      const double r_point = m.px[2 * (size_t)cc], z_point = m.px[2 * (size_t)cc + 1];
      const size_t p1 = c0 + c1 - 1, p2 = c0 + c2 - 1;
      const double r_point1 = m.px[2 * p1], z_point1 = m.px[2 * p1 + 1];
      const double r_point2 = m.px[2 * p2], z_point2 = m.px[2 * p2 + 1];
      const double r_edge1 = half * (r_point + r_point1), z_edge1 = half * (z_point + z_point1);
      const double r_edge2 = half * (r_point + r_point2), z_edge2 = half * (z_point + z_point2);
      g.RadiusFP[2 * (size_t)cc] = half * (r_point + r_edge2);
      g.RadiusFP[2 * (size_t)cc + 1] = half * (r_point + r_edge1);
      g.RadiusEZ[2 * (size_t)cc] = half * (r_zone + r_edge1);
      g.RadiusEZ[2 * (size_t)cc + 1] = half * (r_zone + r_edge2);
      // The reference compares the local corner number with the global
      // neighbour (geometryUCBrz.F90:102,113); kept so every face gets the
      // same one of its two (negated) computations.
      if (c < cfp1) {
        AFP(0, 0, cc) = half * (z_point2 - z_point);
        AFP(1, 0, cc) = half * (r_point - r_point2);
        if (cfp1 <= C) {
          AFP(0, 1, cfp1 - 1) = -AFP(0, 0, cc);
          AFP(1, 1, cfp1 - 1) = -AFP(1, 0, cc);
        }
      }
      if (c < cfp2) {
        AFP(0, 1, cc) = half * (z_point - z_point1);
        AFP(1, 1, cc) = half * (r_point1 - r_point);
        if (cfp2 <= C) {
          AFP(0, 0, cfp2 - 1) = -AFP(0, 1, cc);
          AFP(1, 0, cfp2 - 1) = -AFP(1, 1, cc);
        }
      }
      if (c < c1) {
        AEZ(0, 0, cc) = z_edge1 - z_zone;
        AEZ(1, 0, cc) = r_zone - r_edge1;
        AEZ(0, 1, (int)p1) = -AEZ(0, 0, cc);
        AEZ(1, 1, (int)p1) = -AEZ(1, 0, cc);
      }
      if (c < c2) {
        AEZ(0, 1, cc) = z_zone - z_edge2;
        AEZ(1, 1, cc) = r_edge2 - r_zone;
        AEZ(0, 0, (int)p2) = -AEZ(0, 1, cc);
        AEZ(1, 0, (int)p2) = -AEZ(1, 1, cc);
      }
      const double area1 = std::fabs((r_edge2 - r_point) * (z_zone - z_point) -
                                     (z_edge2 - z_point) * (r_zone - r_point));
      const double area2 = std::fabs((r_zone - r_point) * (z_edge1 - z_point) -
                                     (z_zone - z_point) * (r_edge1 - r_point));
      g.Area[cc] = half * (area1 + area2);
      const double rbar1 = third * (r_point + r_edge2 + r_zone);
      const double rbar2 = third * (r_point + r_edge1 + r_zone);
      g.Volume[cc] = half * (rbar1 * area1 + rbar2 * area2);
    }
    double vz = 0.0;
    for (int c = 0; c < nCorner; ++c) vz += g.Volume[c0 + c];
    g.VolumeZone[z] = vz;
  }
}

// rt/quadrz.F90:184-268: per polar level (last table entry first) a -xi and
// a +xi level, each a zero-weight starting direction (mu = -sin theta), the
// negative-mu directions, the positive-mu directions and a zero-weight
// finishing direction; weights normalised to 1/wtiso = 2 pi and the
// starting/finishing flags set as rt/rtquad.F90:80-123 does; then the
// angular coefficients of rt/AngleCoef2D.F90:30-100 and
// mods/AngleSet_mod.F90:336-347.
void rz_product_quadrature(int npolar, int nazimuthal, RzQuadrature& q) {
  if (npolar < 1 || npolar > 4 || nazimuthal < 1 || nazimuthal > 4)
    throw std::invalid_argument("RZ product quadrature: npolar and nazimuthal must be in 1..4");
  q = RzQuadrature();
  q.npolar = npolar;
  q.nazimuthal = nazimuthal;
  const int A = npolar * 2 * (2 * nazimuthal + 2);
  q.nangles = A;
  q.omega.assign(2 * (size_t)A, 0.0);
  q.weight.assign(A, 0.0);
  const int th1 = kFirst[npolar - 1], ph1 = kFirst[nazimuthal - 1];
  int m = 0;
  auto put = [&](double mu, double xi, double w) {
    q.omega[2 * (size_t)m] = mu;
    q.omega[2 * (size_t)m + 1] = xi;
    q.weight[m] = w;
    ++m;
  };
  for (int jt = npolar - 1; jt >= 0; --jt) {
    const double cosineTheta = kCosTheta[th1 + jt];
    const double sineTheta = std::sqrt(1.0 - cosineTheta * cosineTheta);
    const double xilev = cosineTheta;
    const double wt = kWeightTheta[th1 + jt];
    for (int s = -1; s <= 1; s += 2) {
      const double xi = s * xilev;
      put(-std::sqrt(1.0 - xilev * xilev), xi, 0.0);
      for (int ip = nazimuthal - 1; ip >= 0; --ip)
        put(-sineTheta * kCosPhiRZ[ph1 + ip], xi, wt * kWeightPhiRZ[ph1 + ip]);
      for (int ip = 0; ip < nazimuthal; ++ip)
        put(sineTheta * kCosPhiRZ[ph1 + ip], xi, wt * kWeightPhiRZ[ph1 + ip]);
      put(std::sqrt(1.0 - xilev * xilev), xi, 0.0);
    }
  }
  const double wtiso = 1.0 / (2.0 * M_PI);
  double sumwgt = 0.0;
  for (int a = 0; a < A; ++a) sumwgt += q.weight[a];
  const double fac = 1.0 / (wtiso * sumwgt);
  for (int a = 0; a < A; ++a) q.weight[a] = fac * q.weight[a];
  const double adqtEpsilon = 2.220446049250313e-16;
  q.starting.assign(A, 0);
  q.finishing.assign(A, 0);
  int iang = -1;
  for (int a = 0; a < A; ++a)
    if (q.weight[a] < adqtEpsilon) {
      if (iang == -1) q.starting[a] = 1;
      else q.finishing[a] = 1;
      iang = -iang;
    }
  // Angular coefficients per xi level.
  std::vector<double> alpha(A, 0.0), tau(A, 0.0);
  q.level.assign(A, 0);
  q.nlevels = 0;
  int a = 0;
  while (a < A) {
    if (!q.starting[a]) throw std::logic_error("RZ quadrature: level without a starting direction");
    int e = a + 1;
    while (e < A && !q.starting[e]) ++e;
    double weightLevel = 0.0;
    for (int k = a; k < e; ++k) weightLevel += q.weight[k];
    double Phimh = M_PI, Mumh = q.omega[2 * (size_t)a];
    for (int k = a; k < e; ++k) {
      q.level[k] = q.nlevels;
      if (q.starting[k] || q.finishing[k]) continue;
      alpha[k] = alpha[k - 1] - q.weight[k] * q.omega[2 * (size_t)k];
      const double Phiph = Phimh - q.weight[k] * M_PI / weightLevel;
      const double xi = q.omega[2 * (size_t)k + 1];
      const double Muph = std::sqrt(1.0 - xi * xi) * std::cos(Phiph);
      const double mu = q.omega[2 * (size_t)k];
      if (mu < Mumh || mu > Muph) throw std::logic_error("RZ quadrature: mu not between limits");
      tau[k] = (mu - Mumh) / (Muph - Mumh);
      Phimh = Phiph;
      Mumh = Muph;
    }
    ++q.nlevels;
    a = e;
  }
  q.angDerivFac.assign(A, 0.0);
  q.quadTauW1.assign(A, 1.0);
  q.quadTauW2.assign(A, 0.0);
  for (int k = 0; k < A; ++k)
    if (!q.starting[k] && !q.finishing[k]) {
      q.angDerivFac[k] = q.omega[2 * (size_t)k] + alpha[k] / (q.weight[k] * tau[k]);
      q.quadTauW1[k] = 1.0 / tau[k];
      q.quadTauW2[k] = (1.0 - tau[k]) / tau[k];
    }
}

namespace {

inline double dot2(const double* o, const double* a) { return o[0] * a[0] + o[1] * a[1]; }

// One direction: snneed (2D), findseeds, getDownStreamData, snnext.
int rz_schedule_angle(const RzMesh& m, const RzGeometry& g, const double* omega, int angle,
                      Schedule& s, std::vector<int>& planes, std::string& err) {
  const int Z = m.nzones;
  std::vector<int> needZ(Z, 0);
  std::vector<char> exitFace(kRzFaces * (size_t)Z, 0);
  for (int z = 0; z < Z; ++z) {
    const int c0 = m.cOffSet[z];
    for (int c = 0; c < m.numCorner[z]; ++c) {
      const size_t cc = c0 + c;
      const int face = m.CToFace[2 * cc] - 1;
      const int zo = m.zoneOpp[kRzFaces * (size_t)z + face];
      if (z + 1 < zo) {
        const int fo = m.faceOpp[kRzFaces * (size_t)z + face] - 1;
        const double afpm = dot2(omega, &g.A_fp[4 * cc]);
        if (afpm < 0.0) {
          needZ[z] += 1;
          exitFace[kRzFaces * (size_t)(zo - 1) + fo] = 1;
        } else if (afpm > 0.0) {
          needZ[zo - 1] += 1;
          exitFace[kRzFaces * (size_t)z + face] = 1;
        }
      }
    }
  }
  std::vector<int> listZone;
  listZone.reserve(Z);
  for (int z = 0; z < Z; ++z)
    if (needZ[z] == 0) listZone.push_back(z);
  char buf[160];
  if (listZone.empty()) {
    snprintf(buf, sizeof buf, "angle %d: no seed zone (a zone cycle)", angle + 1);
    err = buf;
    return 3;
  }
  for (int z = 0; z < Z; ++z) {
    const int c0 = m.cOffSet[z], nCorner = m.numCorner[z];
    int need[kRzCorner] = {0}, nDSC[kRzCorner] = {0}, ds[kRzCorner][kRzCF];
    for (int c = 0; c < nCorner; ++c) {
      const size_t cc = c0 + c;
      for (int cf = 0; cf < kRzCF; ++cf) {
        const int cez = m.cEZ[2 * cc + cf];
        if (cez > c + 1) {
          const double aez = dot2(omega, &g.A_ez[4 * cc + 2 * cf]);
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
      if (need[c] != 0) {
        snprintf(buf, sizeof buf, "angle %d: zone %d has a corner dependency cycle", angle + 1,
                 z + 1);
        err = buf;
        return 4;
      }
      nextC[i] = c + 1;
      for (int k = 0; k < nDSC[c]; ++k) need[ds[c][k]] -= 1;
      need[c] = 99;
    }
  }
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
        if (!exitFace[kRzFaces * (size_t)z + f]) continue;
        const int zx = m.zoneOpp[kRzFaces * (size_t)z + f];
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
      snprintf(buf, sizeof buf, "angle %d: zone dependency cycle after %d zones", angle + 1,
               lastZone);
      err = buf;
      return 6;
    }
    newZones = added;
  }
  return 0;
}

}  // namespace

int build_rz_schedule(const RzMesh& m, const RzGeometry& g, const RzQuadrature& q, Schedule& s,
                      std::string& err) {
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
    if (q.finishing[a]) continue;  // finishing directions are not swept (rt/rtorder.F90:54-58)
    std::string e;
    int r = rz_schedule_angle(m, g, &q.omega[2 * (size_t)a], a, s, planes[a], e);
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

// Longest-chain levels over (zone, direction) pairs. Within a direction
// zones are visited in nextZ order (upwind zones first); a direction's
// chain continues its xi level's previous direction at the same zone.
void build_rz_levels(const RzMesh& m, const RzGeometry& g, const RzQuadrature& q,
                     const Schedule& s, RzLevels& lv) {
  const int Z = m.nzones, A = q.nangles;
  std::vector<int> lev((size_t)A * Z, -1);
  int maxLev = -1;
  for (int a = 0; a < A; ++a) {
    if (q.finishing[a]) continue;
    const double* om = &q.omega[2 * (size_t)a];
    const int* nextZ = &s.nextZ[(size_t)a * Z];
    const bool chain = !q.starting[a];
    for (int i = 0; i < Z; ++i) {
      const int z = std::abs(nextZ[i]) - 1;
      int L = chain ? lev[(size_t)(a - 1) * Z + z] + 1 : 0;
      for (int c = 4 * z; c < 4 * z + 4; ++c)
        for (int f = 0; f < kRzCF; ++f) {
          const int up = m.cFP[2 * (size_t)c + f] - 1;
          if (up >= m.ncorner || dot2(om, &g.A_fp[4 * (size_t)c + 2 * f]) >= 0.0) continue;
          L = std::max(L, lev[(size_t)a * Z + up / 4] + 1);
        }
      lev[(size_t)a * Z + z] = L;
      maxLev = std::max(maxLev, L);
    }
  }
  std::vector<std::vector<std::array<int, 2>>> byLevel(maxLev + 1);
  for (int a = 0; a < A; ++a) {
    if (q.finishing[a]) continue;
    const int* nextZ = &s.nextZ[(size_t)a * Z];
    for (int i = 0; i < Z; ++i) {
      const int z = std::abs(nextZ[i]) - 1;
      byLevel[lev[(size_t)a * Z + z]].push_back({z, a});
    }
  }
  lv.items.clear();
  lv.levelOffset.assign(1, 0);
  for (auto& l : byLevel) {
    lv.items.insert(lv.items.end(), l.begin(), l.end());
    lv.levelOffset.push_back((int)lv.items.size());
  }
}

void build_rz_tiles(const RzMesh& m, const RzGeometry& g, const RzQuadrature& q,
                    const Schedule& s, int T, RzTiles& t) {
  t = RzTiles();
  t.T = T;
  const int Z = m.nzones, A = q.nangles;
  if (Z == 0 || T < 1) return;
  // Tiles: zone centroids binned on a grid of about T x T zones per cell.
  double rmin = 1e300, rmax = -1e300, zmin = 1e300, zmax = -1e300;
  for (size_t i = 0; i + 1 < m.points.size(); i += 2) {
    rmin = std::min(rmin, m.points[i]);
    rmax = std::max(rmax, m.points[i]);
    zmin = std::min(zmin, m.points[i + 1]);
    zmax = std::max(zmax, m.points[i + 1]);
  }
  const double bin = T * std::sqrt((rmax - rmin) * (zmax - zmin) / Z);
  if (!(bin > 0.0) || !std::isfinite(bin)) return;
  const int nbr = std::max(1, (int)std::ceil((rmax - rmin) / bin));
  const int nbz = std::max(1, (int)std::ceil((zmax - zmin) / bin));
  std::vector<int> tileOfBin((size_t)nbr * nbz, -1), tileOf(Z);
  std::vector<std::vector<int>> members;
  for (int z = 0; z < Z; ++z) {
    double cr = 0.0, cz = 0.0;
    for (int c = 4 * z; c < 4 * z + 4; ++c) {
      cr += 0.25 * m.px[2 * (size_t)c];
      cz += 0.25 * m.px[2 * (size_t)c + 1];
    }
    const int br = std::min(nbr - 1, std::max(0, (int)std::floor((cr - rmin) / bin)));
    const int bz = std::min(nbz - 1, std::max(0, (int)std::floor((cz - zmin) / bin)));
    int& tid = tileOfBin[(size_t)bz * nbr + br];
    if (tid < 0) {
      tid = (int)members.size();
      members.emplace_back();
    }
    members[tid].push_back(z);
    tileOf[z] = tid;
  }
  const int NT = (int)members.size();
  int maxZ = 0;
  for (auto& mz : members) maxZ = std::max(maxZ, (int)mz.size());
  if (maxZ > kRzMaxTileZones) return;
  // Per swept direction: zone steps inside tiles, the tile graph (Kahn's
  // order; a leftover tile means a cycle) and the tile levels.
  std::vector<unsigned char> zstep((size_t)A * Z, 255);
  std::vector<int> tileLev((size_t)A * NT, -1);
  std::vector<std::vector<int>> succ(NT);
  std::vector<int> indeg(NT), order;
  int maxLev = -1;
  for (int a = 0; a < A; ++a) {
    if (q.finishing[a]) continue;
    const double* om = &q.omega[2 * (size_t)a];
    const int* nextZ = &s.nextZ[(size_t)a * Z];
    unsigned char* st = &zstep[(size_t)a * Z];
    for (auto& v : succ) v.clear();
    std::fill(indeg.begin(), indeg.end(), 0);
    for (int i = 0; i < Z; ++i) {
      const int z = std::abs(nextZ[i]) - 1;
      int k = 0;
      for (int c = 4 * z; c < 4 * z + 4; ++c)
        for (int f = 0; f < kRzCF; ++f) {
          const int up = m.cFP[2 * (size_t)c + f] - 1;
          if (up >= m.ncorner || dot2(om, &g.A_fp[4 * (size_t)c + 2 * f]) >= 0.0) continue;
          const int uz = up / 4;
          if (tileOf[uz] == tileOf[z]) k = std::max(k, st[uz] + 1);
          else succ[tileOf[uz]].push_back(tileOf[z]);
        }
      if (k > 254) return;
      st[z] = (unsigned char)k;
    }
    for (auto& v : succ) {
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      for (int w : v) ++indeg[w];
    }
    order.clear();
    for (int tt = 0; tt < NT; ++tt)
      if (indeg[tt] == 0) order.push_back(tt);
    for (size_t k = 0; k < order.size(); ++k)
      for (int w : succ[order[k]])
        if (--indeg[w] == 0) order.push_back(w);
    if ((int)order.size() != NT) return;
    const bool chain = !q.starting[a];
    int* L = &tileLev[(size_t)a * NT];
    for (int tt : order) {
      if (chain) L[tt] = std::max(L[tt], tileLev[(size_t)(a - 1) * NT + tt] + 1);
      L[tt] = std::max(L[tt], 0);
      for (int w : succ[tt]) L[w] = std::max(L[w], L[tt] + 1);
      maxLev = std::max(maxLev, L[tt]);
    }
  }
  t.ntiles = NT;
  t.maxZones = maxZ;
  t.tileZone.assign((size_t)NT * maxZ, -1);
  for (int tt = 0; tt < NT; ++tt)
    for (size_t i = 0; i < members[tt].size(); ++i)
      t.tileZone[(size_t)tt * maxZ + i] = members[tt][i];
  std::vector<std::vector<std::array<int, 2>>> byLevel(maxLev + 1);
  for (int a = 0; a < A; ++a) {
    if (q.finishing[a]) continue;
    for (int tt = 0; tt < NT; ++tt) byLevel[tileLev[(size_t)a * NT + tt]].push_back({tt, a});
  }
  t.levelOffset.assign(1, 0);
  std::vector<std::pair<int, int>> bySlot;
  for (auto& l : byLevel) {
    for (auto& it : l) {
      const int tt = it[0], a = it[1];
      bySlot.clear();
      for (int i = 0; i < maxZ; ++i) {
        const int z = t.tileZone[(size_t)tt * maxZ + i];
        if (z >= 0) bySlot.push_back({zstep[(size_t)a * Z + z], i});
      }
      std::sort(bySlot.begin(), bySlot.end());
      const int n = bySlot.empty() ? 0 : bySlot.back().first + 1;
      for (int i = 0; i < maxZ; ++i)
        t.order.push_back(i < (int)bySlot.size() ? (unsigned char)bySlot[i].second : (unsigned char)255);
      for (int k = 0, j = 0; k <= maxZ; ++k) {
        while (j < (int)bySlot.size() && bySlot[j].first < k) ++j;
        t.stepOff.push_back((unsigned char)j);
      }
      t.nsteps.push_back(n);
      t.items.push_back(it);
    }
    t.levelOffset.push_back((int)t.items.size());
  }
  t.ok = true;
}

}  // namespace umt
