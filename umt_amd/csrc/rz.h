// Host-side description of a 2D cylindrical (RZ) problem for the UCB RZ
// sweep: quad zones with four corners, two corner faces per corner, in the
// Teton layout (column-major shapes, 1-based index contents, groups fastest).
//   Geometry arrays: mods/Geometry_mod.F90:23-48 (ndim = 2)
//   Angle arrays:    mods/AngleSet_mod.F90:49-73 (2D angular coefficients)
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "host.h"

namespace umt {

constexpr int kRzCF = 2;      // corner faces per quad corner
constexpr int kRzCorner = 4;  // corners per quad zone
constexpr int kRzFaces = 4;   // faces (edges) per quad zone

struct RzMesh {
  int nzones = 0, ncorner = 0, nbelem = 0, npoints = 0;
  std::vector<int> numCorner, cOffSet, zoneFaces;  // (Z)
  std::vector<int> cFP, cEZ, CToFace;              // (2, C)
  std::vector<int> zoneOpp, faceOpp;               // (4, Z)
  std::vector<double> px;                          // (2, C) corner (r, z)
  std::vector<double> points;                      // (2, npoints)
  std::vector<std::array<int, 4>> quads;           // clockwise node ids per zone
};

struct RzGeometry {
  std::vector<double> A_fp, A_ez;          // (2, 2, C)
  std::vector<double> RadiusFP, RadiusEZ;  // (2, C)
  std::vector<double> Area, Volume;        // (C)
  std::vector<double> VolumeZone;          // (Z)
};

struct RzQuadrature {
  int npolar = 0, nazimuthal = 0, nangles = 0, nlevels = 0;
  std::vector<double> omega;   // (2, A): (mu, xi)
  std::vector<double> weight;  // (A), sum = 2 pi
  std::vector<int> starting, finishing;  // (A)
  std::vector<double> angDerivFac, quadTauW1, quadTauW2;  // (A)
  std::vector<int> level;      // (A) 0-based xi level of each direction
};

// Device sweep levels of a whole RZ sweep: item (zone, angle) runs after
// (zone, previous angle of its xi level), whose half-angle flux it reads,
// and after its upwind zones in its own direction. Level L holds every item
// whose longest such chain has L predecessors.
struct RzLevels {
  std::vector<std::array<int, 2>> items;  // (zone, angle), 0-based, level by level
  std::vector<int> levelOffset;           // (levels + 1)
};

// The same sweep in tiles: zones are binned by centroid into tiles of about
// T x T zones; item (tile, angle) runs after the items of its upwind tiles
// and after (tile, previous angle of its xi level). One wave sweeps an
// item's zones in steps, step s holding the zones whose longest chain of
// upwind zones inside the tile has s links, with a barrier between steps,
// so one launch covers several zone levels. ok is false (and the zone
// levels are used) when some direction's tile graph has a cycle or a tile
// holds more than kRzMaxTileZones zones.
constexpr int kRzMaxTileZones = 64;
struct RzTiles {
  bool ok = false;
  int T = 0, maxZones = 0, ntiles = 0;
  std::vector<int> tileZone;              // (ntiles * maxZones) zone or -1
  std::vector<std::array<int, 2>> items;  // (tile, angle), level by level
  std::vector<int> levelOffset;           // (levels + 1)
  std::vector<unsigned char> order;       // (items * maxZones) zone slots by step, 255 = none
  std::vector<unsigned char> stepOff;     // (items * (maxZones + 1)) first entry of each step
  std::vector<int> nsteps;                // (items)
};

// Any quad mesh: points (2, npoints) as (r, z), four node ids per zone
// (0-based). Zones listed counterclockwise are reversed to the clockwise
// order the corner-face conventions need.
void build_rz_mesh(int npoints, const std::vector<double>& points,
                   const std::vector<std::array<int, 4>>& quads, RzMesh& m);
// nr x nz zones on [r0, r1] x [z0, z1], interior nodes moved by up to
// jitter times the cell size (counter-based draws keyed by seed).
void build_rz_structured(int nr, int nz, double r0, double r1, double z0, double z1,
                         double jitter, uint64_t seed, RzMesh& m);
void compute_rz_geometry(const RzMesh& m, RzGeometry& g);
void rz_product_quadrature(int npolar, int nazimuthal, RzQuadrature& q);
// Returns 0, or nonzero with a message (a schedule with cycles).
int build_rz_schedule(const RzMesh& m, const RzGeometry& g, const RzQuadrature& q, Schedule& s,
                      std::string& err);
void build_rz_levels(const RzMesh& m, const RzGeometry& g, const RzQuadrature& q,
                     const Schedule& s, RzLevels& lv);
void build_rz_tiles(const RzMesh& m, const RzGeometry& g, const RzQuadrature& q,
                    const Schedule& s, int T, RzTiles& t);

}  // namespace umt
