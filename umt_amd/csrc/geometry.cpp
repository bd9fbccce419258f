// Corner geometry for the UCB method in xyz geometry, quadrature and the
// counter-based draws used for synthetic cross sections.
//
// compute_geometry follows rt/geometryUCBxyz.F90:63-181: each corner face is
// the sum of two half-side triangles; FP corner-face areas are made equal
// and opposite across interior faces (:116-136), EZ areas equal and opposite
// within a zone (:167-179); corner volumes are pyramids from the zone centre
// (:156-158). Zone and face centres are vertex means
// (mods/Geometry_mod.F90:344-441).
#include <cmath>
#include <stdexcept>

#include "host.h"
#include "qrs45.h"

namespace umt {

void compute_geometry(const Mesh& m, Geometry& g) {
  const int Z = m.nzones, C = m.ncorner;
  g.A_fp.assign(3 * kMaxCF * (size_t)C, 0.0);
  g.A_ez.assign(3 * kMaxCF * (size_t)C, 0.0);
  g.Volume.assign(C, 0.0);
  g.VolumeZone.assign(Z, 0.0);
  auto AFP = [&](int f, int c) { return &g.A_fp[3 * (kMaxCF * (size_t)c + f)]; };
  auto AEZ = [&](int f, int c) { return &g.A_ez[3 * (kMaxCF * (size_t)c + f)]; };
  auto PX = [&](int c) { return &m.px[3 * (size_t)c]; };

  for (int z = 0; z < Z; ++z) {
    const int nCorner = m.numCorner[z], c0 = m.cOffSet[z], nFaces = m.zoneFaces[z];
    double zc[3] = {0, 0, 0};
    for (int c = 0; c < nCorner; ++c)
      for (int d = 0; d < 3; ++d) zc[d] += PX(c0 + c)[d];
    for (int d = 0; d < 3; ++d) zc[d] /= (double)nCorner;
    double fc[kMaxFaces][3] = {};
    int ncf[kMaxFaces] = {0};
    for (int c = 0; c < nCorner; ++c)
      for (int cf = 0; cf < m.nCFaces[c0 + c]; ++cf) {
        int face = m.CToFace[kMaxCF * (size_t)(c0 + c) + cf] - 1;
        ncf[face] += 1;
        for (int d = 0; d < 3; ++d) fc[face][d] += PX(c0 + c)[d];
      }
    for (int f = 0; f < nFaces; ++f)
      for (int d = 0; d < 3; ++d) fc[f][d] = ncf[f] ? fc[f][d] / (double)ncf[f] : 0.0;

    for (int c = 0; c < nCorner; ++c) {
      const int cc = c0 + c;
      const int nCF = m.nCFaces[cc];
      for (int cf1 = 0; cf1 < nCF; ++cf1) {
        const int cf2 = (cf1 + 1) % nCF;
        const int cfp = m.cFP[kMaxCF * (size_t)cc + cf1];  // 1-based
        const int cez1 = m.cEZ[kMaxCF * (size_t)cc + cf1];
        const int cez2 = m.cEZ[kMaxCF * (size_t)cc + cf2];
        const int face = m.CToFace[kMaxCF * (size_t)cc + cf1] - 1;
        const double* p0 = PX(cc);
        const double* p1 = PX(c0 + cez1 - 1);
        const double* p2 = PX(c0 + cez2 - 1);
        double tdl[3], tfl[3], tzl[3], tfz[3], tfe1[3], tfe2[3];
        for (int d = 0; d < 3; ++d) {
          tdl[d] = 0.5 * (p1[d] - p2[d]);
          tfl[d] = fc[face][d] - p0[d];
          tzl[d] = zc[d] - p0[d];
          tfz[d] = fc[face][d] - zc[d];
          tfe1[d] = fc[face][d] - 0.5 * (p1[d] + p0[d]);
          tfe2[d] = fc[face][d] - 0.5 * (p2[d] + p0[d]);
        }
        double afep[3] = {0.5 * (tfl[2] * tdl[1] - tfl[1] * tdl[2]),
                          0.5 * (tfl[0] * tdl[2] - tfl[2] * tdl[0]),
                          0.5 * (tfl[1] * tdl[0] - tfl[0] * tdl[1])};
        const int zoneOpp = m.zoneOpp[kMaxFaces * (size_t)z + face];
        if (zoneOpp > 0) {
          if (zoneOpp > z + 1) {
            for (int d = 0; d < 3; ++d) AFP(cf1, cc)[d] = afep[d];
            const int nb = cfp - 1;
            for (int cf = 0; cf < m.nCFaces[nb]; ++cf)
              if (m.cFP[kMaxCF * (size_t)nb + cf] == cc + 1)
                for (int d = 0; d < 3; ++d) AFP(cf, nb)[d] = -afep[d];
          } else {
            for (int d = 0; d < 3; ++d) afep[d] = AFP(cf1, cc)[d];
          }
        } else if (zoneOpp < 0) {
          for (int d = 0; d < 3; ++d) AFP(cf1, cc)[d] = afep[d];
        }
        double a1[3] = {0.5 * (tfz[2] * tfe1[1] - tfz[1] * tfe1[2]),
                        0.5 * (tfz[0] * tfe1[2] - tfz[2] * tfe1[0]),
                        0.5 * (tfz[1] * tfe1[0] - tfz[0] * tfe1[1])};
        double a2[3] = {0.5 * (tfz[1] * tfe2[2] - tfz[2] * tfe2[1]),
                        0.5 * (tfz[2] * tfe2[0] - tfz[0] * tfe2[2]),
                        0.5 * (tfz[0] * tfe2[1] - tfz[1] * tfe2[0])};
        for (int d = 0; d < 3; ++d) {
          AEZ(cf1, cc)[d] += a1[d];
          AEZ(cf2, cc)[d] += a2[d];
        }
        g.Volume[cc] += (1.0 / 3.0) * std::fabs(tzl[0] * afep[0] + tzl[1] * afep[1] + tzl[2] * afep[2]);
      }
    }
    for (int c = 0; c < nCorner; ++c) {
      const int cc = c0 + c;
      for (int cf1 = 0; cf1 < m.nCFaces[cc]; ++cf1) {
        const int cez = m.cEZ[kMaxCF * (size_t)cc + cf1];
        if (cez > c + 1) {
          const int cn = c0 + cez - 1;
          for (int cf2 = 0; cf2 < m.nCFaces[cn]; ++cf2)
            if (m.cEZ[kMaxCF * (size_t)cn + cf2] == c + 1)
              for (int d = 0; d < 3; ++d) AEZ(cf2, cn)[d] = -AEZ(cf1, cc)[d];
        }
      }
    }
    double vz = 0.0;
    for (int c = 0; c < nCorner; ++c) vz += g.Volume[c0 + c];
    g.VolumeZone[z] = vz;
  }
}

// Product quadrature from the Spence QRS45 polar/azimuthal tables
// (mods/QuadratureData_mod.F90:1144-2857, N = 1..4 only), assembled as
// rt/quadProduct.F90:73-182 does with the polar axis along x (UMT's test
// program's default, driver/test_driver.cc:1198), then normalised so that the weights
// sum to 4*pi (rt/rtquad.F90:95-105).

void product_quadrature(int npolar, int nazimuthal, Quadrature& q) {
  if (npolar < 1 || npolar > 4 || nazimuthal < 1 || nazimuthal > 4)
    throw std::invalid_argument("product quadrature: npolar and nazimuthal must be in 1..4");
  q.npolar = npolar;
  q.nazimuthal = nazimuthal;
  const int nangoct = npolar * nazimuthal;
  q.nangles = 8 * nangoct;
  std::vector<double> ox, oy, oz, w;
  const int phi1 = kFirst[nazimuthal - 1], phi2 = phi1 + nazimuthal - 1;
  const int th1 = kFirst[npolar - 1], th2 = th1 + npolar - 1;
  for (int ip = phi1; ip <= phi2; ++ip) {
    const double cphi = kCosPhi[ip];
    for (int jt = th2; jt >= th1; --jt) {
      const double ct = kCosTheta[jt];
      const double st = std::sqrt(1.0 - ct * ct);
      const double x = ct, y = st * cphi;
      ox.push_back(x);
      oy.push_back(y);
      oz.push_back(std::sqrt(1.0 - x * x - y * y));
      w.push_back(kWeightTheta[jt] * kWeightPhi[ip]);
    }
  }
  static const int sgn[8][3] = {{1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1},
                                {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, -1}};
  q.omega.assign(3 * (size_t)q.nangles, 0.0);
  q.weight.assign(q.nangles, 0.0);
  int nn = 0;
  for (int i = 0; i < nangoct; ++i) {
    for (int o = 0; o < 8; ++o) {
      q.omega[3 * (nn + o) + 0] = sgn[o][0] * ox[i];
      q.omega[3 * (nn + o) + 1] = sgn[o][1] * oy[i];
      q.omega[3 * (nn + o) + 2] = sgn[o][2] * oz[i];
      q.weight[nn + o] = w[i];
    }
    nn += 8;
  }
  double sum = 0.0;
  for (double x : q.weight) sum += x;
  const double wtiso = 1.0 / (4.0 * M_PI);
  const double fac = 1.0 / (wtiso * sum);
  for (double& x : q.weight) x *= fac;
}

// Level-symmetric S2 as rt/quadxyz.F90 builds it for order 2: one direction
// per octant with all three cosines equal to the tabulated 0.577350269189625
// (mods/QuadratureData_mod.F90:711-713), octants numbered as in quadProduct,
// weights normalised to 4*pi (rt/rtquad.F90:95-105). The grey-acceleration
// sweeps use it (gtaorder = 2, driver/test_driver.cc:1206).
void level_symmetric_s2(Quadrature& q) {
  static const int sgn[8][3] = {{1, 1, 1}, {-1, 1, 1}, {-1, -1, 1}, {1, -1, 1},
                                {1, 1, -1}, {-1, 1, -1}, {-1, -1, -1}, {1, -1, -1}};
  const double a = 0.577350269189625;
  q.npolar = 1;
  q.nazimuthal = 1;
  q.nangles = 8;
  q.omega.assign(24, 0.0);
  q.weight.assign(8, 0.0);
  for (int o = 0; o < 8; ++o) {
    for (int d = 0; d < 3; ++d) q.omega[3 * o + d] = sgn[o][d] * a;
    q.weight[o] = 0.5 * M_PI;  // pi/2 per octant, normalised below
  }
  double sum = 0.0;
  for (double w : q.weight) sum += w;
  const double fac = 1.0 / (sum / (4.0 * M_PI));
  for (double& w : q.weight) w *= fac;
}

namespace {
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
}  // namespace

double draw_uniform(uint64_t seed, uint64_t zone, uint64_t group, uint64_t field) {
  uint64_t h = splitmix64(seed);
  h = splitmix64(h ^ zone);
  h = splitmix64(h ^ ((group << 8) | field));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace umt
