// C ABI of the RZ sweep (include/umt_rz.h): problems built on the host
// (rz.cpp) and the device-resident RZ context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/umt_rz.h"
#include "common.h"
#include "kernels.h"
#include "rz.h"

using umt::dalloc;
using umt::dfree;
using umt::fail;

struct umt_rz_problem {
  umt::RzMesh m;
  umt::RzGeometry g;
  umt::RzQuadrature q;
  umt::Schedule s;
  umt::RzLevels lv;
  umt::RzTiles tiles;
};

struct umt_rz_ctx {
  int Z = 0, C = 0, nb = 0, G = 0, A = 0, nlev = 0, precision = 64, K = 1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void *Afp = nullptr, *Aez = nullptr, *Rfp = nullptr, *Rez = nullptr, *area = nullptr,
       *vol = nullptr;
  int* cFP = nullptr;
  double *vol64 = nullptr, *volZ = nullptr;
  void *S = nullptr, *sigt = nullptr;
  double *sigs = nullptr, *q = nullptr, *phi = nullptr, *phibar = nullptr, *phibarOld = nullptr;
  unsigned long long* bits = nullptr;
  umt::RzAngle* angles = nullptr;
  double* weight = nullptr;
  int2* items = nullptr;
  std::vector<int> levelOffset;
  // Tile levels (umt::RzTiles), used when tilesOk and useTiles (off by
  // default: each step of a tile pays the whole dependent-load chain of one
  // item, which the zone levels hide behind their many independent items).
  bool tilesOk = false, useTiles = false;
  int tileMaxZones = 0;
  int2* tileItems = nullptr;
  unsigned char *tileOrder = nullptr, *tileStepOff = nullptr;
  int *tileNsteps = nullptr, *tileZone = nullptr;
  std::vector<int> tileLevelOffset;
  void* psi = nullptr;   // (G, C + nb, A)
  void* psiM = nullptr;  // (G, C, xi levels)
  std::vector<hipEvent_t> evLevel;
  bool haveSigt = false, haveScattering = false;
  std::vector<double> sigtHost;  // sigma_t as set; the sweep sees sigma_t + tau
  std::vector<double> sigsHost;  // sigma_s as set (checked against sigma_t)
  double tau = 0.0;              // 1/(c dt): time-dependent sweeps when > 0
  void* psiPrev = nullptr;       // previous cycle's psi (G, C, A), sweep precision
  // Captured sweeps: the level launches and the phi reduction (sweep), or
  // the level launches alone (levels, for timing), per groups-per-lane K.
  bool useGraphs = true;
  struct Graph {
    int K;
    bool levelsOnly;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  size_t rsize() const { return precision == 32 ? sizeof(float) : sizeof(double); }
  long stride() const { return (long)(C + nb) * G; }
  bool tiled() const { return tilesOk && useTiles; }
  int levels() const {
    return (int)(tiled() ? tileLevelOffset.size() : levelOffset.size()) - 1;
  }
};

namespace {

// Tiles of about kRzTile x kRzTile zones per workgroup.
constexpr int kRzTile = 4;

int finish_rz_problem(umt_rz_problem* p, int npolar, int nazimuthal) {
  try {
    umt::compute_rz_geometry(p->m, p->g);
    umt::rz_product_quadrature(npolar, nazimuthal, p->q);
    std::string err;
    if (int rc = umt::build_rz_schedule(p->m, p->g, p->q, p->s, err))
      return fail(60 + rc, "RZ schedule: %s", err.c_str());
    umt::build_rz_levels(p->m, p->g, p->q, p->s, p->lv);
    umt::build_rz_tiles(p->m, p->g, p->q, p->s, kRzTile, p->tiles);
  } catch (const std::exception& e) {
    return fail(60, "RZ problem: %s", e.what());
  }
  return 0;
}

template <typename T>
void copy_out(const std::vector<T>& v, T* dst) {
  if (dst) std::copy(v.begin(), v.end(), dst);
}

int upload(umt_rz_ctx* x, void* dst, const double* src, size_t n) {
  if (n == 0) return 0;
  if (x->precision == 64) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, x->stream));
  } else {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, x->stream));
  }
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

// The sweep's total cross section: sigma_t + tau (control/setTotalOpacity.F90).
int upload_sigt(umt_rz_ctx* x) {
  std::vector<double> st(x->sigtHost);
  for (double& v : st) v += x->tau;
  return upload(x, x->sigt, st.data(), st.size());
}

template <typename Real>
int sweep_impl(umt_rz_ctx* x, bool timeLevels, bool levelsOnly = false) {
  umt::RzLevelArgs<Real> a{};
  a.G = x->G;
  a.C = x->C;
  a.Afp = (const Real*)x->Afp;
  a.Aez = (const Real*)x->Aez;
  a.Rfp = (const Real*)x->Rfp;
  a.Rez = (const Real*)x->Rez;
  a.area = (const Real*)x->area;
  a.vol = (const Real*)x->vol;
  a.cFP = x->cFP;
  a.S = (const Real*)x->S;
  a.sigt = (const Real*)x->sigt;
  a.angles = x->angles;
  a.psi = (Real*)x->psi;
  a.psiAngleStride = x->stride();
  a.psiM = (Real*)x->psiM;
  a.tauPsi = x->tau > 0.0 ? (const Real*)x->psiPrev : nullptr;
  a.tau = (Real)x->tau;
  if (x->tiled()) {
    umt::RzTileArgs<Real> t{};
    t.p = a;
    t.tileZone = x->tileZone;
    t.maxZones = x->tileMaxZones;
    const int blocks = x->G / x->K;
    t.gbPerChunk = blocks >= 4 ? 4 : (blocks >= 2 ? 2 : 1);
    for (int h = 0; h < x->levels(); ++h) {
      const int i0 = x->tileLevelOffset[h], n = x->tileLevelOffset[h + 1] - i0;
      t.items = x->tileItems + i0;
      t.order = x->tileOrder + (size_t)i0 * x->tileMaxZones;
      t.stepOff = x->tileStepOff + (size_t)i0 * (x->tileMaxZones + 1);
      t.nsteps = x->tileNsteps + i0;
      if (timeLevels) HIP_TRY(hipEventRecord(x->evLevel[2 * h], x->stream));
      HIP_TRY(umt::launch_rz_sweep_tile<Real>(t, n, x->K, x->stream));
      if (timeLevels) HIP_TRY(hipEventRecord(x->evLevel[2 * h + 1], x->stream));
    }
  } else
  for (int h = 0; h < x->levels(); ++h) {
    a.items = x->items + x->levelOffset[h];
    a.nItems = x->levelOffset[h + 1] - x->levelOffset[h];
    if (timeLevels) HIP_TRY(hipEventRecord(x->evLevel[2 * h], x->stream));
    HIP_TRY(umt::launch_rz_sweep_level<Real>(a, x->K, x->stream));
    if (timeLevels) HIP_TRY(hipEventRecord(x->evLevel[2 * h + 1], x->stream));
  }
  if (!levelsOnly)
    HIP_TRY(umt::launch_phi_from_psi_acc<Real>((const Real*)x->psi, x->weight, x->A, x->stride(),
                                               (long)x->C * x->G, x->phi, false, x->stream));
  return 0;
}

int launch_eager(umt_rz_ctx* x, bool timeLevels, bool levelsOnly) {
  return x->precision == 32 ? sweep_impl<float>(x, timeLevels, levelsOnly)
                            : sweep_impl<double>(x, timeLevels, levelsOnly);
}

void drop_graphs(umt_rz_ctx* x) {
  for (auto& g : x->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  x->graphs.clear();
}

// The sweep's launches as a graph (kernel nodes only), built on first use
// for the current groups per lane and replayed afterwards; the arrays it
// reads and writes stay at fixed addresses for the context's lifetime.
int sweep_graph(umt_rz_ctx* x, bool levelsOnly, hipGraphExec_t* out) {
  for (auto& g : x->graphs)
    if (g.K == x->K && g.levelsOnly == levelsOnly) {
      *out = g.exec;
      return 0;
    }
  umt_rz_ctx::Graph g{x->K, levelsOnly, nullptr, nullptr};
  HIP_TRY(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  if (int rc = launch_eager(x, false, levelsOnly)) {
    hipGraph_t dummy = nullptr;
    (void)hipStreamEndCapture(x->stream, &dummy);
    if (dummy) (void)hipGraphDestroy(dummy);
    return rc;
  }
  HIP_TRY(hipStreamEndCapture(x->stream, &g.graph));
  HIP_TRY(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  x->graphs.push_back(g);
  *out = g.exec;
  return 0;
}

int sweep(umt_rz_ctx* x, bool timeLevels = false, bool levelsOnly = false) {
  if (!x->haveSigt) return fail(30, "umt_rz_ctx_set_sigt first");
  if (x->useGraphs && !timeLevels) {
    hipGraphExec_t exec = nullptr;
    if (int rc = sweep_graph(x, levelsOnly, &exec)) return rc;
    HIP_TRY(hipGraphLaunch(exec, x->stream));
    return 0;
  }
  return launch_eager(x, timeLevels, levelsOnly);
}

int create_impl(const umt_rz_problem* p, int G, int precision, umt_rz_ctx* x) {
  if (G < 1) return fail(1, "groups must be positive");
  if (precision != 32 && precision != 64) return fail(1, "precision must be 32 or 64");
  const umt::RzMesh& m = p->m;
  const umt::RzGeometry& g = p->g;
  const umt::RzQuadrature& q = p->q;
  x->Z = m.nzones;
  x->C = m.ncorner;
  x->nb = m.nbelem;
  x->G = G;
  x->A = q.nangles;
  x->nlev = q.nlevels;
  x->precision = precision;
  x->K = G % 4 == 0 ? 4 : (G % 2 == 0 ? 2 : 1);
  HIP_TRY(hipStreamCreate(&x->stream));
  HIP_TRY(hipEventCreate(&x->ev0));
  HIP_TRY(hipEventCreate(&x->ev1));
  const size_t rs = x->rsize(), C = x->C, Z = x->Z, Gs = G, A = x->A;
  const size_t nItems = p->lv.items.size();
  int rc;
  if ((rc = dalloc((char**)&x->Afp, 4 * C * rs)) || (rc = dalloc((char**)&x->Aez, 4 * C * rs)) ||
      (rc = dalloc((char**)&x->Rfp, 2 * C * rs)) || (rc = dalloc((char**)&x->Rez, 2 * C * rs)) ||
      (rc = dalloc((char**)&x->area, C * rs)) || (rc = dalloc((char**)&x->vol, C * rs)) ||
      (rc = dalloc(&x->cFP, 2 * C)) || (rc = dalloc(&x->vol64, C)) || (rc = dalloc(&x->volZ, Z)) ||
      (rc = dalloc((char**)&x->S, Gs * C * rs)) || (rc = dalloc((char**)&x->sigt, Gs * Z * rs)) ||
      (rc = dalloc(&x->sigs, Gs * Z)) || (rc = dalloc(&x->q, Gs * Z)) ||
      (rc = dalloc(&x->phi, Gs * C)) || (rc = dalloc(&x->phibar, Gs * Z)) ||
      (rc = dalloc(&x->phibarOld, Gs * Z)) || (rc = dalloc(&x->bits, 2)) ||
      (rc = dalloc(&x->angles, A)) || (rc = dalloc(&x->weight, A)) ||
      (rc = dalloc(&x->items, nItems)) ||
      (rc = dalloc((char**)&x->psi, (size_t)x->stride() * A * rs)) ||
      (rc = dalloc((char**)&x->psiM, Gs * C * (size_t)x->nlev * rs)))
    return rc;
  if ((rc = upload(x, x->Afp, g.A_fp.data(), 4 * C)) || (rc = upload(x, x->Aez, g.A_ez.data(), 4 * C)) ||
      (rc = upload(x, x->Rfp, g.RadiusFP.data(), 2 * C)) ||
      (rc = upload(x, x->Rez, g.RadiusEZ.data(), 2 * C)) ||
      (rc = upload(x, x->area, g.Area.data(), C)) || (rc = upload(x, x->vol, g.Volume.data(), C)))
    return rc;
  HIP_TRY(hipMemcpyAsync(x->cFP, m.cFP.data(), 2 * C * sizeof(int), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(x->vol64, g.Volume.data(), C * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(x->volZ, g.VolumeZone.data(), Z * sizeof(double), hipMemcpyHostToDevice, x->stream));
  std::vector<umt::RzAngle> an(A);
  for (size_t k = 0; k < A; ++k) {
    an[k].mu = q.omega[2 * k];
    an[k].xi = q.omega[2 * k + 1];
    an[k].fac = q.angDerivFac[k];
    an[k].w1 = q.quadTauW1[k];
    an[k].w2 = q.quadTauW2[k];
    an[k].level = q.level[k];
    an[k].flags = (q.starting[k] ? 1 : 0) | ((k + 1 < A && q.finishing[k + 1]) ? 2 : 0);
  }
  HIP_TRY(hipMemcpyAsync(x->angles, an.data(), A * sizeof(umt::RzAngle), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(x->weight, q.weight.data(), A * sizeof(double), hipMemcpyHostToDevice, x->stream));
  std::vector<int2> items(nItems);
  // XCD-aware order: workgroup w of a level launch runs on XCD w mod 8, so
  // slot w takes the next workgroup's worth of items (in the level's own
  // order) whose zone lies in zone range w mod 8 of eight equal zone-id
  // ranges, or from the fullest range once its own is used up. A zone and
  // its upwind neighbours then mostly meet in one XCD's L2 from level to
  // level. Items of a level are independent, so the order changes no result.
  {
    constexpr int kXcd = 8;
    const int perWg = std::max(1, 64 / std::max(1, x->G / x->K));
    const std::vector<int>& lo = p->lv.levelOffset;
    std::array<std::vector<int2>, kXcd> part;
    for (size_t h = 0; h + 1 < lo.size(); ++h) {
      for (auto& v : part) v.clear();
      for (int i = lo[h]; i < lo[h + 1]; ++i) {
        const int z = p->lv.items[i][0];
        part[(int)((long)z * kXcd / x->Z)].push_back(make_int2(z, p->lv.items[i][1]));
      }
      std::array<size_t, kXcd> next{};
      size_t o = lo[h];
      for (long slot = 0; o < (size_t)lo[h + 1]; ++slot) {
        int r = (int)(slot % kXcd);
        if (next[r] == part[r].size()) {
          int best = -1;
          for (int t = 0; t < kXcd; ++t)
            if (next[t] < part[t].size() &&
                (best < 0 || part[t].size() - next[t] > part[best].size() - next[best]))
              best = t;
          r = best;
        }
        const size_t take = std::min((size_t)perWg, part[r].size() - next[r]);
        for (size_t k = 0; k < take; ++k) items[o++] = part[r][next[r]++];
      }
    }
  }
  HIP_TRY(hipMemcpyAsync(x->items, items.data(), nItems * sizeof(int2), hipMemcpyHostToDevice, x->stream));
  x->levelOffset = p->lv.levelOffset;
  const umt::RzTiles& tl = p->tiles;
  x->tilesOk = tl.ok;
  if (tl.ok) {
    x->tileMaxZones = tl.maxZones;
    x->tileLevelOffset = tl.levelOffset;
    const size_t ni = tl.items.size();
    if ((rc = dalloc(&x->tileItems, ni)) || (rc = dalloc(&x->tileOrder, tl.order.size())) ||
        (rc = dalloc(&x->tileStepOff, tl.stepOff.size())) ||
        (rc = dalloc(&x->tileNsteps, ni)) || (rc = dalloc(&x->tileZone, tl.tileZone.size())))
      return rc;
    std::vector<int2> ti(ni);
    for (size_t i = 0; i < ni; ++i) ti[i] = make_int2(tl.items[i][0], tl.items[i][1]);
    HIP_TRY(hipMemcpyAsync(x->tileItems, ti.data(), ni * sizeof(int2), hipMemcpyHostToDevice, x->stream));
    HIP_TRY(hipMemcpyAsync(x->tileOrder, tl.order.data(), tl.order.size(), hipMemcpyHostToDevice, x->stream));
    HIP_TRY(hipMemcpyAsync(x->tileStepOff, tl.stepOff.data(), tl.stepOff.size(), hipMemcpyHostToDevice,
                           x->stream));
    HIP_TRY(hipMemcpyAsync(x->tileNsteps, tl.nsteps.data(), ni * sizeof(int), hipMemcpyHostToDevice, x->stream));
    HIP_TRY(hipMemcpyAsync(x->tileZone, tl.tileZone.data(), tl.tileZone.size() * sizeof(int),
                           hipMemcpyHostToDevice, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  }
  const long nPsi = x->stride() * (long)A, nPm = (long)Gs * C * x->nlev;
  if (precision == 32) {
    HIP_TRY(umt::launch_fill_zero<float>((float*)x->psi, nPsi, x->stream));
    HIP_TRY(umt::launch_fill_zero<float>((float*)x->psiM, nPm, x->stream));
    HIP_TRY(umt::launch_fill_zero<float>((float*)x->S, (long)Gs * C, x->stream));
  } else {
    HIP_TRY(umt::launch_fill_zero<double>((double*)x->psi, nPsi, x->stream));
    HIP_TRY(umt::launch_fill_zero<double>((double*)x->psiM, nPm, x->stream));
    HIP_TRY(umt::launch_fill_zero<double>((double*)x->S, (long)Gs * C, x->stream));
  }
  HIP_TRY(umt::launch_fill_zero<double>(x->phi, (long)Gs * C, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>(x->phibarOld, (long)Gs * Z, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>(x->sigs, (long)Gs * Z, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>(x->q, (long)Gs * Z, x->stream));
  x->evLevel.resize(2 * std::max(x->levelOffset.size(), x->tileLevelOffset.size()));
  for (auto& e : x->evLevel) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

}  // namespace

extern "C" {

int umt_rz_problem_structured(int nr, int nz, double r0, double r1, double z0, double z1,
                              double jitter, uint64_t seed, int npolar, int nazimuthal,
                              umt_rz_problem** out) {
  if (!out) return fail(1, "null output pointer");
  *out = nullptr;
  umt_rz_problem* p = new umt_rz_problem();
  try {
    umt::build_rz_structured(nr, nz, r0, r1, z0, z1, jitter, seed, p->m);
  } catch (const std::exception& e) {
    delete p;
    return fail(60, "RZ mesh: %s", e.what());
  }
  if (int rc = finish_rz_problem(p, npolar, nazimuthal)) {
    delete p;
    return rc;
  }
  *out = p;
  return 0;
}

int umt_rz_problem_from_quad_mesh(int npoints, const double* rz, int nzones, const int* quads,
                                  int npolar, int nazimuthal, umt_rz_problem** out) {
  if (!out || !rz || !quads) return fail(1, "null argument");
  *out = nullptr;
  if (npoints < 4 || nzones < 1) return fail(1, "bad sizes");
  for (long i = 0; i < 2L * npoints; ++i)
    if (!std::isfinite(rz[i])) return fail(17, "coordinate %ld is not finite", i + 1);
  for (long i = 0; i < npoints; ++i)
    if (rz[2 * i] < 0.0) return fail(17, "point %ld has r < 0", i + 1);
  std::vector<std::array<int, 4>> q((size_t)nzones);
  for (long z = 0; z < nzones; ++z)
    for (int k = 0; k < 4; ++k) {
      const int id = quads[4 * z + k];
      if (id < 0 || id >= npoints) return fail(18, "zone %ld node %d: id %d out of range", z + 1, k + 1, id);
      q[z][k] = id;
    }
  umt_rz_problem* p = new umt_rz_problem();
  try {
    umt::build_rz_mesh(npoints, std::vector<double>(rz, rz + 2L * npoints), q, p->m);
  } catch (const std::exception& e) {
    delete p;
    return fail(19, "RZ mesh: %s", e.what());
  }
  if (int rc = finish_rz_problem(p, npolar, nazimuthal)) {
    delete p;
    return rc;
  }
  *out = p;
  return 0;
}

void umt_rz_problem_destroy(umt_rz_problem* p) { delete p; }

int umt_rz_problem_dims(const umt_rz_problem* p, int64_t* d) {
  if (!p || !d) return fail(1, "null argument");
  d[0] = p->m.nzones;
  d[1] = p->m.ncorner;
  d[2] = p->m.nbelem;
  d[3] = p->q.nangles;
  d[4] = p->m.npoints;
  d[5] = (int64_t)p->s.zonesInPlane.size();
  d[6] = p->q.nlevels;
  d[7] = (int64_t)p->lv.levelOffset.size() - 1;
  return 0;
}

int umt_rz_problem_mesh(const umt_rz_problem* p, int* numCorner, int* cOffSet, int* cFP, int* cEZ,
                        int* CToFace, int* zoneOpp, int* faceOpp, double* px) {
  if (!p) return fail(1, "null problem");
  copy_out(p->m.numCorner, numCorner);
  copy_out(p->m.cOffSet, cOffSet);
  copy_out(p->m.cFP, cFP);
  copy_out(p->m.cEZ, cEZ);
  copy_out(p->m.CToFace, CToFace);
  copy_out(p->m.zoneOpp, zoneOpp);
  copy_out(p->m.faceOpp, faceOpp);
  copy_out(p->m.px, px);
  return 0;
}

int umt_rz_problem_points(const umt_rz_problem* p, double* rz, int* quads) {
  if (!p) return fail(1, "null problem");
  copy_out(p->m.points, rz);
  if (quads)
    for (size_t z = 0; z < p->m.quads.size(); ++z)
      for (int k = 0; k < 4; ++k) quads[4 * z + k] = p->m.quads[z][k];
  return 0;
}

int umt_rz_problem_geometry(const umt_rz_problem* p, double* A_fp, double* A_ez, double* RadiusFP,
                            double* RadiusEZ, double* Area, double* Volume, double* VolumeZone) {
  if (!p) return fail(1, "null problem");
  copy_out(p->g.A_fp, A_fp);
  copy_out(p->g.A_ez, A_ez);
  copy_out(p->g.RadiusFP, RadiusFP);
  copy_out(p->g.RadiusEZ, RadiusEZ);
  copy_out(p->g.Area, Area);
  copy_out(p->g.Volume, Volume);
  copy_out(p->g.VolumeZone, VolumeZone);
  return 0;
}

int umt_rz_problem_quadrature(const umt_rz_problem* p, double* omega, double* weight, int* starting,
                              int* finishing, double* angDerivFac, double* quadTauW1,
                              double* quadTauW2) {
  if (!p) return fail(1, "null problem");
  copy_out(p->q.omega, omega);
  copy_out(p->q.weight, weight);
  copy_out(p->q.starting, starting);
  copy_out(p->q.finishing, finishing);
  copy_out(p->q.angDerivFac, angDerivFac);
  copy_out(p->q.quadTauW1, quadTauW1);
  copy_out(p->q.quadTauW2, quadTauW2);
  return 0;
}

int umt_rz_problem_schedule(const umt_rz_problem* p, int* nHyperPlanes, int* zonesInPlane,
                            int* nextZ, int* nextC) {
  if (!p) return fail(1, "null problem");
  copy_out(p->s.nHyperPlanes, nHyperPlanes);
  copy_out(p->s.zonesInPlane, zonesInPlane);
  copy_out(p->s.nextZ, nextZ);
  copy_out(p->s.nextC, nextC);
  return 0;
}

void umt_rz_ctx_destroy(umt_rz_ctx* x) {
  if (!x) return;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  drop_graphs(x);
  dfree(x->Afp); dfree(x->Aez); dfree(x->Rfp); dfree(x->Rez); dfree(x->area); dfree(x->vol);
  dfree(x->cFP); dfree(x->vol64); dfree(x->volZ); dfree(x->S); dfree(x->sigt); dfree(x->sigs);
  dfree(x->q); dfree(x->phi); dfree(x->phibar); dfree(x->phibarOld); dfree(x->bits);
  dfree(x->angles); dfree(x->weight); dfree(x->items); dfree(x->psi); dfree(x->psiM);
  dfree(x->psiPrev);
  dfree(x->tileItems); dfree(x->tileOrder); dfree(x->tileStepOff); dfree(x->tileNsteps);
  dfree(x->tileZone);
  for (auto e : x->evLevel) (void)hipEventDestroy(e);
  if (x->ev0) (void)hipEventDestroy(x->ev0);
  if (x->ev1) (void)hipEventDestroy(x->ev1);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
}

int umt_rz_ctx_create(const umt_rz_problem* p, int groups, int precision, umt_rz_ctx** out) {
  if (!out || !p) return fail(1, "null argument");
  *out = nullptr;
  umt_rz_ctx* x = new umt_rz_ctx();
  if (int rc = create_impl(p, groups, precision, x)) {
    umt_rz_ctx_destroy(x);
    return rc;
  }
  *out = x;
  return 0;
}

int umt_rz_ctx_dims(const umt_rz_ctx* x, int64_t* d) {
  if (!x || !d) return fail(1, "null argument");
  d[0] = x->Z; d[1] = x->C; d[2] = x->nb; d[3] = x->G; d[4] = x->A;
  d[5] = x->levels();
  d[6] = x->levelOffset.back();
  d[7] = x->tiled() ? 1 : 0;
  return 0;
}

int umt_rz_ctx_set_sigt(umt_rz_ctx* x, const double* sigt) {
  if (!x || !sigt) return fail(1, "null argument");
  const long n = (long)x->G * x->Z;
  for (long i = 0; i < n; ++i)
    if (!(sigt[i] > 0.0) || !std::isfinite(sigt[i])) return fail(30, "Sigt entry %ld is not positive and finite", i + 1);
  x->sigtHost.assign(sigt, sigt + n);
  if (int rc = upload_sigt(x)) return rc;
  x->haveSigt = true;
  return 0;
}

int umt_rz_ctx_set_source(umt_rz_ctx* x, const double* S) {
  if (!x || !S) return fail(1, "null argument");
  const long n = (long)x->G * x->C;
  for (long i = 0; i < n; ++i)
    if (!std::isfinite(S[i])) return fail(31, "STotal entry %ld is not finite", i + 1);
  return upload(x, x->S, S, n);
}

int umt_rz_ctx_set_scattering(umt_rz_ctx* x, const double* sigs, const double* q) {
  if (!x || !sigs || !q) return fail(1, "null argument");
  const long n = (long)x->G * x->Z;
  for (long i = 0; i < n; ++i)
    if (!(sigs[i] >= 0.0) || !std::isfinite(sigs[i]) || !std::isfinite(q[i]))
      return fail(32, "sigs/q entry %ld is negative or not finite", i + 1);
  HIP_TRY(hipMemcpyAsync(x->sigs, sigs, n * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(x->q, q, n * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  x->haveScattering = true;
  x->sigsHost.assign(sigs, sigs + n);
  return 0;
}

int umt_rz_ctx_sweep(umt_rz_ctx* x) {
  if (!x) return fail(1, "null context");
  if (int rc = sweep(x)) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_rz_ctx_get_phi(const umt_rz_ctx* x, double* phi) {
  if (!x || !phi) return fail(1, "null argument");
  HIP_TRY(hipMemcpyAsync(phi, x->phi, (size_t)x->G * x->C * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_rz_ctx_set_phi(umt_rz_ctx* x, const double* phi) {
  if (!x || !phi) return fail(1, "null argument");
  HIP_TRY(hipMemcpyAsync(x->phi, phi, (size_t)x->G * x->C * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_rz_ctx_get_psi(const umt_rz_ctx* x, double* psi) {
  if (!x || !psi) return fail(1, "null argument");
  const size_t n = (size_t)x->stride() * x->A;
  if (x->precision == 64) {
    HIP_TRY(hipMemcpyAsync(psi, x->psi, n * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  } else {
    std::vector<float> f(n);
    HIP_TRY(hipMemcpyAsync(f.data(), x->psi, n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    for (size_t i = 0; i < n; ++i) psi[i] = f[i];
  }
  return 0;
}

int umt_rz_ctx_source_iteration(umt_rz_ctx* x, double tol, int maxIters, int* iters,
                                double* residual, double* hist) {
  if (!x || maxIters < 0) return fail(1, "bad argument");
  if (!x->haveScattering) return fail(33, "umt_rz_ctx_set_scattering first");
  for (size_t i = 0; i < x->sigsHost.size() && i < x->sigtHost.size(); ++i)
    if (x->sigsHost[i] > x->sigtHost[i])
      return fail(52, "sigs entry %zu exceeds Sigt (scattering ratio above 1)", i + 1);
  const long nGC = (long)x->G * x->C, nGZ = (long)x->G * x->Z;
  // The old zone averages come from the starting phi.
  HIP_TRY(umt::launch_rz_residual(x->phi, x->vol64, x->volZ, x->G, nGZ, x->phibar, x->phibarOld,
                                  1e-6, x->bits, x->stream));
  int it = 0;
  double res = 1.0;
  while (it < maxIters) {
    if (x->precision == 32)
      HIP_TRY(umt::launch_rz_source_update<float>(x->phi, x->sigs, x->q, x->G, nGC, (float*)x->S, x->stream));
    else
      HIP_TRY(umt::launch_rz_source_update<double>(x->phi, x->sigs, x->q, x->G, nGC, (double*)x->S, x->stream));
    if (int rc = sweep(x)) return rc;
    HIP_TRY(umt::launch_rz_residual(x->phi, x->vol64, x->volZ, x->G, nGZ, x->phibar, x->phibarOld,
                                    1e-6, x->bits, x->stream));
    unsigned long long b[2];
    HIP_TRY(hipMemcpyAsync(b, x->bits, sizeof b, hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    double r;
    std::memcpy(&r, &b[1], sizeof r);
    if (!std::isfinite(r)) return fail(34, "RZ source iteration residual is not finite at iteration %d", it + 1);
    res = r;
    if (hist) hist[it] = res;
    ++it;
    if (res < tol) break;
  }
  if (iters) *iters = it;
  if (residual) *residual = res;
  return res < tol ? 0 : 1;
}

int umt_rz_ctx_set_tau(umt_rz_ctx* x, double tau) {
  if (!x) return fail(1, "null context");
  if (!(tau >= 0.0) || !std::isfinite(tau)) return fail(48, "tau must be finite and >= 0");
  HIP_TRY(hipStreamSynchronize(x->stream));  // no replay in flight when the graphs go
  const long n = (long)x->G * x->C * x->A;
  if (tau > 0.0 && !x->psiPrev) {
    if (int rc = dalloc((char**)&x->psiPrev, (size_t)n * x->rsize())) return rc;
    if (x->precision == 32) HIP_TRY(umt::launch_fill_zero<float>((float*)x->psiPrev, n, x->stream));
    else HIP_TRY(umt::launch_fill_zero<double>((double*)x->psiPrev, n, x->stream));
  }
  x->tau = tau;
  drop_graphs(x);  // the level launches carry tau and the Psi pointer
  if (x->haveSigt)
    if (int rc = upload_sigt(x)) return rc;
  return 0;
}

int umt_rz_ctx_set_psi_prev(umt_rz_ctx* x, const double* Psi) {
  if (!x || !Psi) return fail(1, "null argument");
  if (!x->psiPrev) return fail(49, "umt_rz_ctx_set_tau with tau > 0 first");
  const long n = (long)x->G * x->C * x->A;
  for (long i = 0; i < n; ++i)
    if (!std::isfinite(Psi[i])) return fail(31, "Psi entry %ld is not finite", i + 1);
  return upload(x, x->psiPrev, Psi, n);
}

int umt_rz_ctx_get_psi_prev(const umt_rz_ctx* x, double* Psi) {
  if (!x || !Psi) return fail(1, "null argument");
  if (!x->psiPrev) return fail(49, "umt_rz_ctx_set_tau with tau > 0 first");
  const size_t n = (size_t)x->G * x->C * x->A;
  if (x->precision == 64) {
    HIP_TRY(hipMemcpyAsync(Psi, x->psiPrev, n * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  } else {
    std::vector<float> f(n);
    HIP_TRY(hipMemcpyAsync(f.data(), x->psiPrev, n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    for (size_t i = 0; i < n; ++i) Psi[i] = f[i];
  }
  return 0;
}

int umt_rz_ctx_save_psi(umt_rz_ctx* x) {
  if (!x) return fail(1, "null context");
  if (!x->psiPrev) return fail(49, "umt_rz_ctx_set_tau with tau > 0 first");
  if (!x->haveScattering) return fail(33, "umt_rz_ctx_save_psi needs umt_rz_ctx_set_scattering first");
  // The cycle's final sweep (rt/rtmainsn.F90:238-246): source from the
  // current phi, one sweep reading the previous cycle's Psi, then Psi = the
  // psi of every direction (SweepUCBrz.F90:268-276).
  const long nGC = (long)x->G * x->C;
  if (x->precision == 32)
    HIP_TRY(umt::launch_rz_source_update<float>(x->phi, x->sigs, x->q, x->G, nGC, (float*)x->S, x->stream));
  else
    HIP_TRY(umt::launch_rz_source_update<double>(x->phi, x->sigs, x->q, x->G, nGC, (double*)x->S, x->stream));
  if (int rc = sweep(x)) return rc;
  const size_t rs = x->rsize();
  HIP_TRY(hipMemcpy2DAsync(x->psiPrev, nGC * rs, x->psi, (size_t)x->stride() * rs, nGC * rs,
                           (size_t)x->A, hipMemcpyDeviceToDevice, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_rz_ctx_set_groups_per_lane(umt_rz_ctx* x, int K) {
  if (!x) return fail(1, "null context");
  if (!umt::groups_per_lane_supported(x->G, K))
    return fail(35, "groups per lane %d not supported for %d groups (1, 2 or 4, dividing G)", K, x->G);
  x->K = K;
  return 0;
}

int umt_rz_ctx_autotune(umt_rz_ctx* x, int reps, int* bestK, double* msPerK) {
  if (!x || reps < 1) return fail(1, "bad argument");
  if (!x->haveSigt) return fail(30, "umt_rz_ctx_set_sigt first");
  static const int cand[3] = {1, 2, 4};
  double best = 1e300;
  int bk = x->K;
  for (int i = 0; i < 3; ++i) {
    if (msPerK) msPerK[i] = -1.0;
    if (!umt::groups_per_lane_supported(x->G, cand[i])) continue;
    x->K = cand[i];
    if (int rc = sweep(x)) return rc;  // warm-up (and graph capture)
    HIP_TRY(hipEventRecord(x->ev0, x->stream));
    for (int r = 0; r < reps; ++r)
      if (int rc = sweep(x)) return rc;
    HIP_TRY(hipEventRecord(x->ev1, x->stream));
    HIP_TRY(hipEventSynchronize(x->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, x->ev0, x->ev1));
    const double ms = t / reps;
    if (msPerK) msPerK[i] = ms;
    if (ms < best) {
      best = ms;
      bk = cand[i];
    }
  }
  x->K = bk;
  if (bestK) *bestK = bk;
  return 0;
}

int umt_rz_ctx_set_tiles(umt_rz_ctx* x, int on) {
  if (!x) return fail(1, "null context");
  if (on && !x->tilesOk) return fail(62, "this RZ problem has no tile schedule (a tile graph with a cycle)");
  HIP_TRY(hipStreamSynchronize(x->stream));
  x->useTiles = on != 0;
  drop_graphs(x);
  return 0;
}

int umt_rz_ctx_set_graphs(umt_rz_ctx* x, int on) {
  if (!x) return fail(1, "null context");
  x->useGraphs = on != 0;
  return 0;
}

int umt_rz_ctx_time_sweeps(umt_rz_ctx* x, int warmup, int reps, double* msSweep,
                           double* msLevelKernels) {
  if (!x || warmup < 0 || reps < 1) return fail(1, "bad argument");
  for (int i = 0; i < warmup; ++i)
    if (int rc = sweep(x)) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  float t = 0.f;
  HIP_TRY(hipEventRecord(x->ev0, x->stream));
  for (int r = 0; r < reps; ++r)
    if (int rc = sweep(x)) return rc;
  HIP_TRY(hipEventRecord(x->ev1, x->stream));
  HIP_TRY(hipEventSynchronize(x->ev1));
  HIP_TRY(hipEventElapsedTime(&t, x->ev0, x->ev1));
  if (msSweep) *msSweep = t / reps;
  if (msLevelKernels) {
    double kernelMs = 0.0;
    if (x->useGraphs) {
      // The level launches alone, replayed back to back from a graph.
      if (int rc = sweep(x, false, true)) return rc;
      HIP_TRY(hipEventRecord(x->ev0, x->stream));
      for (int r = 0; r < reps; ++r)
        if (int rc = sweep(x, false, true)) return rc;
      HIP_TRY(hipEventRecord(x->ev1, x->stream));
      HIP_TRY(hipEventSynchronize(x->ev1));
      float lt = 0.f;
      HIP_TRY(hipEventElapsedTime(&lt, x->ev0, x->ev1));
      kernelMs = lt / reps;
    } else {
      if (int rc = sweep(x, true)) return rc;
      HIP_TRY(hipStreamSynchronize(x->stream));
      for (int h = 0; h < x->levels(); ++h) {
        float lt = 0.f;
        HIP_TRY(hipEventElapsedTime(&lt, x->evLevel[2 * h], x->evLevel[2 * h + 1]));
        kernelMs += lt;
      }
    }
    *msLevelKernels = kernelMs;
  }
  return 0;
}

}  // extern "C"
