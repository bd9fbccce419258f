// Host-side problem description for the UCB xyz sweep.
//
// Arrays follow the Teton layout so they can be handed to the drop-in entry
// points unchanged: column-major shapes (first index fastest), index
// *contents* 1-based (corner, zone, face and boundary-element numbers),
// groups fastest in every (G, ...) array.
//   Geometry arrays: mods/Geometry_mod.F90:23-48
//   Schedule arrays: mods/AngleSet_mod.F90:77-102
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace umt {

constexpr int kMaxCF = 3;       // corner faces per hex corner
constexpr int kMaxCorner = 8;   // corners per hex zone
constexpr int kMaxFaces = 6;    // faces per hex zone

struct Mesh {
  int nx = 0, ny = 0, nz = 0;
  int nzones = 0, ncorner = 0, nbelem = 0, npoints = 0;
  std::vector<int> numCorner, cOffSet, zoneFaces;       // (Z)
  std::vector<int> nCFaces;                             // (C)
  std::vector<int> cFP, cEZ, CToFace;                   // (maxcf, C)
  std::vector<int> zoneOpp, faceOpp;                    // (maxFaces, Z)
  std::vector<int> nCPerFace;                           // (maxFaces, Z)
  std::vector<double> px;                               // (3, C) corner coords
  std::vector<int> bdyToC, bdyToZone;                   // (nbelem)
  std::vector<int> cornerPoint;                         // (C) 0-based point id
  std::vector<double> points;                           // (3, npoints)
  std::vector<int> zoneTile;                            // (Z) 0-based tile-order id
};

struct Geometry {
  std::vector<double> A_fp, A_ez;   // (3, maxcf, C)
  std::vector<double> Volume;       // (C)
  std::vector<double> VolumeZone;   // (Z)
};

struct Quadrature {
  int npolar = 0, nazimuthal = 0, nangles = 0;
  std::vector<double> omega;   // (3, A)
  std::vector<double> weight;  // (A), sum = 4*pi
};

struct Schedule {
  // Per angle (1-based contents, as rtorder/snnext produce them).
  int nangles = 0, nzones = 0, ncorner = 0;
  std::vector<int> nHyperPlanes;   // (A)
  std::vector<int> planeOffset;    // (A+1) offsets into zonesInPlane
  std::vector<int> zonesInPlane;   // concatenated per angle
  std::vector<int> nextZ;          // (Z, A) sign = non-cyclic(+)/cyclic(-)
  std::vector<int> nextC;          // (C, A) local corner order per zone
  std::vector<int> numCycles;      // (A) corners on each angle's cycle list
  std::vector<int> cycleList;      // 1-based corners, angle by angle (rt/rtorder.F90:62-90)
  int maxPlanes = 0;
  long totalCycles() const { return (long)cycleList.size(); }
};

void build_tiled_mesh(int nx, int ny, int nz, bool kdtree, Mesh& m);
// Any hex mesh: points (3, npoints), 8 node ids per zone (0-based, Blueprint
// / VTK hexahedron order), optional kd-tree zone order.
void build_hex_mesh(int npoints, const std::vector<double>& points,
                    const std::vector<std::array<int, 8>>& hexPts, bool kdtree, Mesh& m);
void compute_geometry(const Mesh& m, Geometry& g);
void product_quadrature(int npolar, int nazimuthal, Quadrature& q);
// Level-symmetric S2 (8 directions), the grey-acceleration quadrature.
void level_symmetric_s2(Quadrature& q);
// Returns 0 on success, nonzero (and a message) when the reference's
// schedule would stop with a fatal error. Dependency cycles are broken as
// the reference breaks them (Schedule::numCycles, cycleList).
// Corner faces with |Omega.A| <= parallelTol * |A|_1 are treated as parallel
// to Omega (no dependency); parallelTol = 0 is the reference's exact test.
int build_schedule(const Mesh& m, const Geometry& g, const Quadrature& q,
                   Schedule& s, std::string& err, double parallelTol = 0.0);

// Counter-based uniform draw in [0,1) keyed by (seed, zone, group, field).
double draw_uniform(uint64_t seed, uint64_t zone, uint64_t group, uint64_t field);

}  // namespace umt
