// Tiled hexahedral mesh of the unit cube and its Teton corner connectivity.
//
// The tile is the 24-quad pattern UMT's driver builds with Conduit's tiled
// example (README.md:75,79; img/tile.jpg; driver/test_driver.cc:1787-1816):
// a 20x20 square with four corner squares, a central 2x2 block and sixteen
// quadrilaterals cut by the diagonals and the diamond through the edge
// midpoints. Conduit is not vendored, so this generator defines the mesh;
// every consumer (sweep, oracle) receives the same arrays from here.
// Each tile is one hex layer thick in z; nx*ny*nz tiles fill [0,1]^3.
//
// build_hex_mesh takes any hex mesh as explicit points and 8 node ids per
// zone in the order of a Conduit Blueprint / VTK hexahedron (nodes 0-3 one
// face, counter-clockwise seen from nodes 4-7 above them), the form in which
// interface/TetonBlueprint.cc:774-890 receives UMT's meshes, and builds the
// same Teton corner arrays from it.
//
// Corner connectivity follows aux/setTetonZone.F90:143-191: for each corner,
// the corner faces are chained cyclically so that the EZ neighbour of corner
// face i is the counter-clockwise neighbour on zone face i, and face i+1 is
// the face whose CCW neighbour is face i's clockwise neighbour.
#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <unordered_map>

#include "host.h"

namespace umt {
namespace {

// Base quadrant (x,y in [0,10] of the 20x20 tile), counter-clockwise from +z.
const int kBaseQuads[6][4][2] = {
    {{0, 0}, {3, 0}, {3, 3}, {0, 3}},      // corner square
    {{3, 0}, {10, 0}, {5, 5}, {3, 3}},     // lower quad under the diagonal
    {{10, 0}, {10, 7}, {7, 7}, {5, 5}},    // quad between diamond and centre
    {{0, 3}, {3, 3}, {5, 5}, {0, 10}},     // mirror of the lower quad
    {{0, 10}, {5, 5}, {7, 7}, {7, 10}},    // mirror of the centre quad
    {{7, 7}, {10, 7}, {10, 10}, {7, 10}},  // central square
};

// Local hex faces (0-based local corners). Corners 0-3 are the bottom quad
// (counter-clockwise from +z), 4-7 the top quad above them. Each face is wound
// so that its right-hand normal points into the zone; with the corner-face
// area formula of rt/geometryUCBxyz.F90:100-112 this makes A_fp outward.
const int kHexFaces[6][4] = {
    {0, 1, 2, 3},  // bottom (-z)
    {4, 7, 6, 5},  // top (+z)
    {0, 4, 5, 1},  // side over bottom edge 0-1
    {1, 5, 6, 2},  // side over bottom edge 1-2
    {2, 6, 7, 3},  // side over bottom edge 2-3
    {3, 7, 4, 0},  // side over bottom edge 3-0
};

struct Quad2 {
  std::array<int, 4> p;  // 2D point ids
};

void tile_quads(std::vector<std::array<std::array<int, 2>, 4>>& out) {
  out.clear();
  for (int refl = 0; refl < 4; ++refl) {
    const bool fx = refl & 1, fy = refl & 2;
    for (int q = 0; q < 6; ++q) {
      std::array<std::array<int, 2>, 4> v;
      for (int i = 0; i < 4; ++i) {
        int x = kBaseQuads[q][i][0], y = kBaseQuads[q][i][1];
        v[i] = {fx ? 20 - x : x, fy ? 20 - y : y};
      }
      if (fx != fy) std::swap(v[1], v[3]);  // a single reflection flips winding
      out.push_back(v);
    }
  }
}

void kd_order(std::vector<int>& ids, const std::vector<double>& cen, int lo, int hi) {
  if (hi - lo <= 1) return;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int i = lo; i < hi; ++i)
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], cen[3 * ids[i] + d]);
      mx[d] = std::max(mx[d], cen[3 * ids[i] + d]);
    }
  int axis = 0;
  for (int d = 1; d < 3; ++d)
    if (mx[d] - mn[d] > mx[axis] - mn[axis]) axis = d;
  std::sort(ids.begin() + lo, ids.begin() + hi, [&](int a, int b) {
    double ca = cen[3 * a + axis], cb = cen[3 * b + axis];
    return ca < cb || (ca == cb && a < b);
  });
  int mid = lo + (hi - lo) / 2;
  kd_order(ids, cen, lo, mid);
  kd_order(ids, cen, mid, hi);
}

}  // namespace

void build_tiled_mesh(int nx, int ny, int nz, bool kdtree, Mesh& m) {
  if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("tile counts must be >= 1");
  m = Mesh();
  m.nx = nx; m.ny = ny; m.nz = nz;

  std::vector<std::array<std::array<int, 2>, 4>> tq;
  tile_quads(tq);

  // 2D points shared between tiles, keyed by integer coordinates.
  std::unordered_map<int64_t, int> pid;
  std::vector<std::array<int, 2>> p2;
  std::vector<Quad2> quads;
  auto key = [](int64_t X, int64_t Y) { return (X << 24) | Y; };
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      for (auto& q : tq) {
        Quad2 qq;
        for (int v = 0; v < 4; ++v) {
          int X = 20 * i + q[v][0], Y = 20 * j + q[v][1];
          auto it = pid.find(key(X, Y));
          if (it == pid.end()) {
            int id = (int)p2.size();
            pid.emplace(key(X, Y), id);
            p2.push_back({X, Y});
            qq.p[v] = id;
          } else {
            qq.p[v] = it->second;
          }
        }
        quads.push_back(qq);
      }
  const int n2 = (int)p2.size();
  m.npoints = n2 * (nz + 1);
  m.points.resize(3 * (size_t)m.npoints);
  for (int k = 0; k <= nz; ++k)
    for (int p = 0; p < n2; ++p) {
      size_t id = (size_t)k * n2 + p;
      m.points[3 * id + 0] = (double)p2[p][0] / (20.0 * nx);
      m.points[3 * id + 1] = (double)p2[p][1] / (20.0 * ny);
      m.points[3 * id + 2] = (double)k / (double)nz;
    }

  // Hexes in tile order: layer k, tile (j,i), quad.
  const int nq = (int)quads.size();
  const int Z = nq * nz;
  std::vector<std::array<int, 8>> hexPts(Z);
  for (int k = 0; k < nz; ++k)
    for (int q = 0; q < nq; ++q) {
      auto& h = hexPts[(size_t)k * nq + q];
      for (int v = 0; v < 4; ++v) {
        h[v] = k * n2 + quads[q].p[v];
        h[4 + v] = (k + 1) * n2 + quads[q].p[v];
      }
    }
  std::vector<double> points = std::move(m.points);
  const int npoints = m.npoints;
  build_hex_mesh(npoints, points, hexPts, kdtree, m);
  m.nx = nx; m.ny = ny; m.nz = nz;
}

void build_hex_mesh(int npoints, const std::vector<double>& points,
                    const std::vector<std::array<int, 8>>& hexPts, bool kdtree, Mesh& m) {
  const int Z = (int)hexPts.size();
  if (Z < 1) throw std::invalid_argument("mesh: no zones");
  for (const auto& h : hexPts)
    for (int v : h)
      if (v < 0 || v >= npoints) throw std::invalid_argument("mesh: hex node index out of range");
  m = Mesh();
  m.npoints = npoints;
  m.points = points;

  // Zone order: kd-tree over zone centroids (driver/test_driver.cc:247,1813).
  std::vector<int> order(Z);
  std::iota(order.begin(), order.end(), 0);
  if (kdtree) {
    std::vector<double> cen(3 * (size_t)Z, 0.0);
    for (int z = 0; z < Z; ++z)
      for (int v = 0; v < 8; ++v)
        for (int d = 0; d < 3; ++d) cen[3 * z + d] += m.points[3 * (size_t)hexPts[z][v] + d] / 8.0;
    kd_order(order, cen, 0, Z);
  }

  m.nzones = Z;
  m.ncorner = 8 * Z;
  const int C = m.ncorner;
  m.numCorner.assign(Z, 8);
  m.zoneFaces.assign(Z, 6);
  m.cOffSet.resize(Z);
  m.zoneTile.resize(Z);
  m.cornerPoint.resize(C);
  m.px.resize(3 * (size_t)C);
  for (int z = 0; z < Z; ++z) {
    m.cOffSet[z] = 8 * z;
    m.zoneTile[z] = order[z];
    for (int v = 0; v < 8; ++v) {
      int p = hexPts[order[z]][v];
      m.cornerPoint[8 * z + v] = p;
      for (int d = 0; d < 3; ++d) m.px[3 * (size_t)(8 * z + v) + d] = m.points[3 * (size_t)p + d];
    }
  }

  // Face matching: sorted point ids of each face.
  struct FaceRef { int zone, face; };
  std::unordered_map<std::string, FaceRef> faceMap;
  faceMap.reserve((size_t)Z * 4);
  m.zoneOpp.assign(kMaxFaces * (size_t)Z, 0);
  m.faceOpp.assign(kMaxFaces * (size_t)Z, 0);
  m.nCPerFace.assign(kMaxFaces * (size_t)Z, 4);
  auto fkey = [&](int z, int f) {
    std::array<int, 4> s;
    for (int i = 0; i < 4; ++i) s[i] = m.cornerPoint[8 * z + kHexFaces[f][i]];
    std::sort(s.begin(), s.end());
    return std::string(reinterpret_cast<const char*>(s.data()), sizeof(s));
  };
  for (int z = 0; z < Z; ++z)
    for (int f = 0; f < 6; ++f) {
      std::string k = fkey(z, f);
      auto it = faceMap.find(k);
      if (it == faceMap.end()) {
        faceMap.emplace(k, FaceRef{z, f});
        m.zoneOpp[6 * (size_t)z + f] = -1;  // boundary until matched
      } else {
        int zn = it->second.zone, fn = it->second.face;
        m.zoneOpp[6 * (size_t)z + f] = zn + 1;
        m.faceOpp[6 * (size_t)z + f] = fn + 1;
        m.zoneOpp[6 * (size_t)zn + fn] = z + 1;
        m.faceOpp[6 * (size_t)zn + fn] = f + 1;
        faceMap.erase(it);
      }
    }

  // Boundary elements, numbered per (zone, face, corner-in-face) in zone
  // order, as setTetonZone numbers them for a single boundary condition
  // (aux/setTetonZone.F90:89-118).
  std::vector<int> cornerOpp(24 * (size_t)Z, 0);  // (4 corners, 6 faces, Z)
  m.nbelem = 0;
  for (int z = 0; z < Z; ++z)
    for (int f = 0; f < 6; ++f) {
      int zn = m.zoneOpp[6 * (size_t)z + f];
      for (int i = 0; i < 4; ++i) {
        int lc = kHexFaces[f][i];
        int& co = cornerOpp[24 * (size_t)z + 4 * f + i];
        if (zn < 0) {
          ++m.nbelem;
          m.bdyToC.push_back(8 * z + lc + 1);
          m.bdyToZone.push_back(z + 1);
          co = C + m.nbelem;
        } else {
          int p = m.cornerPoint[8 * z + lc];
          int found = -1;
          for (int v = 0; v < 8; ++v)
            if (m.cornerPoint[8 * (zn - 1) + v] == p) found = v;
          if (found < 0) throw std::runtime_error("mesh: unmatched corner across face");
          co = 8 * (zn - 1) + found + 1;
        }
      }
    }

  // Corner-face chaining (aux/setTetonZone.F90:143-191).
  m.nCFaces.assign(C, 0);
  m.cFP.assign(kMaxCF * (size_t)C, 0);
  m.cEZ.assign(kMaxCF * (size_t)C, 0);
  m.CToFace.assign(kMaxCF * (size_t)C, 0);
  for (int z = 0; z < Z; ++z) {
    int cnt[8] = {0};
    int cCW[8][kMaxCF], cCCW[8][kMaxCF], cOpp[8][kMaxCF], fID[8][kMaxCF];
    for (int f = 0; f < 6; ++f) {
      const int numC = 4;
      for (int i = 0; i < numC; ++i) {
        int iCW = (i + 1) % numC, iCCW = (i + numC - 1) % numC;
        int c = kHexFaces[f][i];
        int s = cnt[c]++;
        if (s >= kMaxCF) throw std::runtime_error("mesh: corner with more than 3 faces");
        cCW[c][s] = kHexFaces[f][iCW] + 1;
        cCCW[c][s] = kHexFaces[f][iCCW] + 1;
        cOpp[c][s] = cornerOpp[24 * (size_t)z + 4 * f + i];
        fID[c][s] = f + 1;
      }
    }
    for (int c = 0; c < 8; ++c) {
      size_t cc = 8 * (size_t)z + c;
      int n = cnt[c];
      m.nCFaces[cc] = n;
      m.cEZ[kMaxCF * cc + 0] = cCCW[c][0];
      m.cFP[kMaxCF * cc + 0] = cOpp[c][0];
      m.CToFace[kMaxCF * cc + 0] = fID[c][0];
      int last = cCW[c][0];
      for (int i = 1; i < n; ++i) {
        bool ok = false;
        for (int ii = 1; ii < n; ++ii)
          if (cCCW[c][ii] == last) {
            m.cEZ[kMaxCF * cc + i] = cCCW[c][ii];
            m.cFP[kMaxCF * cc + i] = cOpp[c][ii];
            m.CToFace[kMaxCF * cc + i] = fID[c][ii];
            last = cCW[c][ii];
            ok = true;
            break;
          }
        if (!ok) throw std::runtime_error("mesh: corner faces do not close");
      }
    }
  }
}

}  // namespace umt
