// Host runtime of the MI355X UCB sweep: problem construction, the
// device-resident sweep context and the drop-in entry points of
// include/umt_sweep.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/umt_sweep.h"
#include "common.h"
#include "host.h"
#include "kernels.h"

namespace {
thread_local std::string g_err;
thread_local int g_dropin_status = 0;
}  // namespace

namespace umt {

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Canonical hex corner topology (1-based EZ neighbours per corner-face slot),
// the layout the sweep kernels are specialised for.
static const int kCanonicalEZ[8][3] = {{4, 2, 5}, {1, 3, 6}, {2, 4, 7}, {3, 1, 8},
                                       {6, 8, 1}, {7, 5, 2}, {8, 6, 3}, {5, 7, 4}};

int check_hex_topology(int Z, int C, const int* numCorner, const int* cOffSet,
                       const int* nCFaces, const int* cEZ) {
  if (C != 8 * Z) return fail(10, "ncorner (%d) != 8 * nzones (%d): only hex zones are supported", C, Z);
  for (int z = 0; z < Z; ++z) {
    if (numCorner[z] != 8 || cOffSet[z] != 8 * z)
      return fail(11, "zone %d: %d corners at offset %d (need 8 corners at offset %d)", z + 1,
                  numCorner[z], cOffSet[z], 8 * z);
    for (int c = 0; c < 8; ++c) {
      const long cc = 8L * z + c;
      if (nCFaces[cc] != 3) return fail(12, "corner %ld has %d corner faces (need 3)", cc + 1, nCFaces[cc]);
      for (int f = 0; f < 3; ++f)
        if (cEZ[3 * cc + f] != kCanonicalEZ[c][f])
          return fail(13, "zone %d corner %d: corner-face order differs from the hex layout the kernels use",
                      z + 1, c + 1);
    }
  }
  return 0;
}

}  // namespace umt

using umt::check_hex_topology;
using umt::dalloc;
using umt::dfree;
using umt::fail;

struct umt_problem {
  umt::Mesh m;
  umt::Geometry g;
  umt::Quadrature q;
  umt::Schedule s;
  umt::Quadrature q2;  // level-symmetric S2 of the grey-acceleration sweeps
  umt::Schedule s2;
  std::string gtaError;  // why the S2 schedule is unavailable ("" when it is built)
};

// Grey-acceleration sweeps use directions that lie exactly in some faces
// of the tile (|Omega.A| is rounding noise there); such corner faces are
// treated as parallel below this fraction of the face area.
static const double kGtaParallelTol = 1e-12;
// The FP32 sweep's Omega.A carries ~1e-7 |A|_1 of rounding (FP32 directions
// and areas), so its S2 sweeps treat faces below 1e-5 |A|_1 as parallel.
static const double kGtaParallelTolFP32 = 1e-5;

// One quadrature's sweep: directions, device-resident hyperplane batches
// (level h = plane h of every angle) and psi for all its angles.
// Angles are swept in blocks of anglesPerBlock (all of them by default):
// block b holds angles [b * anglesPerBlock, ...), its own levels (plane h of
// each of its angles) and reuses one psi buffer of anglesPerBlock angles.
struct SweepSet {
  int A = 0;
  int anglesPerBlock = 0;
  int groupsPerLane = 0;     // K of this set's sweeps (0: the context's)
  double parTol = 0.0;       // parallel-face tolerance of schedule and sweep
  double* omega = nullptr;   // (3, A)
  double* weight = nullptr;  // (A)
  int2* items = nullptr;     // (zone, angle), block by block, level by level
  std::vector<int> levelOffset;  // item offsets of the levels, all blocks
  long zoneVisits = 0;           // distinct (level, zone) pairs over all levels
  std::vector<int> blockLevel;   // first level of each block (+ end)
  std::vector<int2> allItems;    // the unblocked levels (plane h of every angle)
  std::vector<int> allLevelOffset;
  void* psi = nullptr;       // (G, C + nbelem, anglesPerBlock), sweep precision
  // Dependency cycles: cyclic zones in the schedule (nextZ < 0), and the
  // corners whose psi is read before it is computed, per angle
  // (AngleSet numCycles / cycleOffSet / cycleList), with their psi from the
  // previous sweep (SetData cyclePsi).
  bool cyclicZones = false;
  std::vector<int> cycleOffset;  // (A + 1), empty without cycle lists
  int2* cycleEntries = nullptr;  // (corner, angle) 0-based, angle by angle
  void* cyclePsi = nullptr;      // (G, total entries), sweep precision
  long totalCycles() const { return cycleOffset.empty() ? 0 : cycleOffset.back(); }
  std::vector<hipEvent_t> evLevel;
  // Captured sweeps, keyed by source, cross section, groups per lane and
  // the phi they write.
  struct Graph {
    const void* S;
    const void* sigt;
    int K;
    const double* phi;
    bool accumulate;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  int blocks() const { return (int)blockLevel.size() - 1; }
  int levels() const { return (int)levelOffset.size() - 1; }
};

// Grey-transport acceleration state (per-group one-group GTA).
struct GtaState {
  bool enabled = false;
  void* sigt2 = nullptr;  // stretched sigma_t / eps, (G, Z), sweep precision
  double *eps = nullptr;  // 1/sqrt(1 - c), (G, Z)
  void* S2 = nullptr;     // grey sweep source, (G, C), sweep precision
  double *F = nullptr, *r = nullptr, *d = nullptr, *Ad = nullptr, *As = nullptr, *P1 = nullptr;
  double *gsrc = nullptr, *phiHalf = nullptr, *pzOld = nullptr;
  double *partial = nullptr, *dots = nullptr, *coef = nullptr, *err = nullptr;
  int dotBlocks = 0, errBlocks = 0;
  // Solver kernel, numbered as the reference driver's GTA kernel option
  // (driver/test_driver.cc:1520-1531): 1 grey + stretched-TSA sweeps, 2 the
  // transfer-matrix solver (Size%useNewGTASolver).
  int solver = 1;
  double* lu = nullptr;       // kernel 2: T, then its LU factors, 8 (G, C) arrays
  bool luValid = false;       // lu holds the factors of the current opacities
  int2* localItems = nullptr; // kernel 2: every (zone, grey direction), zone by zone
  int maxSweeps = 21;         // grey_max_sweeps (mods/default_iter_controls_mod.F90)
  double epsPoint = 1.0e-6;   // grey_reltol
  // GTA%enforceHardGTAIterMax: stop the grey iteration at maxSweeps whatever
  // its error; off for kernel 1 (mods/GreyAcceleration_mod.F90:122), on for
  // kernel 2 (:169).
  bool hardIterMax = false;
  double epsGreyInit = 1.0e-5;  // rt/rtmainsn.F90:129
  long long greyIterations = 0, greySweeps = 0;
};

// The learned correction network: FP32 parameters with Adam moments, their
// packed BF16 MFMA operands, and the training samples (features X, scaled
// sweep change a, scaled target correction b).
struct umt_learner {
  hipStream_t stream = nullptr;
  float* params = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;  // kMlpParams gradient entries + the squared-error sum
  unsigned short* packed = nullptr;
  long long adamSteps = 0;
  float* X = nullptr;
  float* a = nullptr;
  float* b = nullptr;
  long n = 0, cap = 0;
  int* perm = nullptr;
  long permCap = 0;
  float* partial = nullptr;
  long partialCap = 0;  // rows of kMlpParams + 1
  float* lossLog = nullptr;
  long lossCap = 0;
};

// Learned acceleration state of a context.
struct LearnState {
  umt_learner* net = nullptr;
  int every = 2;       // sweeps between corrections
  bool guard = true;   // stop correcting once a corrected step raises the residual
  bool collect = false;
  int collectStride = 1;   // record every collectStride-th eligible state
  int* zoneNbr = nullptr;   // (6, Z) 0-based face neighbours, -1 on the boundary
  double* sigt64 = nullptr; // (G, Z) FP64 sigma_t for the features
  double *pp = nullptr, *d = nullptr, *dp = nullptr, *rel = nullptr, *ratio = nullptr;
  double *ta = nullptr, *tb = nullptr, *tpart = nullptr;
  unsigned long long* maxRel = nullptr;
  unsigned long long* scratchBits = nullptr;  // [0] zone-average max, [1] non-finite correction flag
  float* X = nullptr;
  float* r = nullptr;
  struct Pending {
    float* X;
    double* ph;
    double* d;
    double res;
  };
  std::vector<Pending> pending;
  long long corrections = 0;
  int fellBack = 0;
};

struct umt_ctx {
  int Z = 0, C = 0, nb = 0, G = 0, precision = 64;
  hipStream_t stream = nullptr;
  // Geometry (sweep precision) and connectivity.
  void* Afp = nullptr;
  void* Aez = nullptr;
  void* vol = nullptr;
  int* cFP = nullptr;
  double* vol64 = nullptr;
  double* volZ = nullptr;
  SweepSet sn;  // the S_N quadrature
  SweepSet s2;  // S2 grey sweeps
  // Fields.
  void* S = nullptr;     // (G, C)
  void* sigt = nullptr;  // (G, Z)
  double* sigs = nullptr;
  double* q = nullptr;
  double* phi = nullptr;
  double* phiPrev = nullptr;
  double* phibar = nullptr;
  double* phibarOld = nullptr;
  unsigned long long* bits = nullptr;
  bool haveScattering = false;
  bool haveSigt = false;
  int groupsPerLane = 1;  // K of the sweep kernels
  bool useGraphs = false; // replay captured level launches
  int stallWindow = 0;    // stop source iteration after this many iterations without a new best residual
  int acceleration = 0;   // 0 plain source iteration, 1 GTA, 2 learned correction, 3 both
  std::vector<double> sigtHost, sigsHost;  // last uploaded cross sections (GTA opacities)
  double nIn = 0.0;       // mean incoming interior FP faces per (corner, angle)
  double tau = 0.0;       // 1/(c dt): time-dependent sweeps when > 0
  void* psiPrev = nullptr;  // previous cycle's psi, (G, C + nbelem, A), sweep precision
  void* psiInc = nullptr;   // incident boundary psi, (G, nbelem, A), sweep precision (null: vacuum)
  std::vector<int> cFPHost;  // (3, C) corner-face neighbours, for splitting cyclic planes
  GtaState gta;
  LearnState learn;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  size_t rsize() const { return precision == 32 ? sizeof(float) : sizeof(double); }
  long psiStride() const { return (long)(C + nb) * G; }
};

namespace {

template <typename Real>
umt::SweepLevelArgs<Real> level_args(const umt_ctx* x, const SweepSet& set, const void* S,
                                     const void* sigt, int level, int angle0) {
  umt::SweepLevelArgs<Real> a{};
  a.items = set.items + set.levelOffset[level];
  a.nItems = set.levelOffset[level + 1] - set.levelOffset[level];
  a.G = x->G;
  a.groupsPerLane = set.groupsPerLane ? set.groupsPerLane : x->groupsPerLane;
  a.C = x->C;
  a.Afp = (const Real*)x->Afp;
  a.Aez = (const Real*)x->Aez;
  a.vol = (const Real*)x->vol;
  a.cFP = x->cFP;
  a.S = (const Real*)S;
  a.sigt = (const Real*)sigt;
  a.omega = set.omega;
  a.psi = (Real*)set.psi;
  a.psiAngleStride = x->psiStride();
  a.angle0 = angle0;
  a.parTol = (Real)set.parTol;
  a.boundaryIn = (&set == &x->sn && x->psiInc) ? 1 : 0;
  if (&set == &x->sn && x->tau > 0.0) {  // tau Psi source of the S_N sweep only
    a.tauPsi = (const Real*)x->psiPrev;
    a.tau = (Real)x->tau;
  }
  return a;
}

// The launches of one sweep: per angle block its levels, then its share of
// phi (FP64, angle order).
template <typename Real>
int launch_sweep(umt_ctx* x, SweepSet& set, const void* S, const void* sigt, double* phiOut,
                 bool accumulate, bool timeLevels, void* savePsi = nullptr) {
  if (set.cyclicZones && set.cycleOffset.empty())
    return fail(15, "the schedule has cyclic zones (nextZ < 0) but no cycle lists: "
                    "umt_ctx_set_cycles first");
  for (int b = 0; b < set.blocks(); ++b) {
    const int a0 = b * set.anglesPerBlock;
    const int nA = std::min(set.anglesPerBlock, set.A - a0);
    long c0 = 0, nc = 0;
    if (!set.cycleOffset.empty()) {
      c0 = set.cycleOffset[a0];
      nc = set.cycleOffset[a0 + nA] - c0;
    }
    // Incident boundary fluxes of the block's angles (SweepUCBxyz.F90:95-97).
    if (&set == &x->sn && x->psiInc)
      HIP_TRY(umt::launch_boundary_in<Real>((const Real*)x->psiInc, (long)x->nb * x->G,
                                            (long)x->C * x->G, a0, nA, x->psiStride(),
                                            (Real*)set.psi, x->stream));
    // initFromCycleList for the block's angles (SetSweep.F90:149)
    HIP_TRY(umt::launch_cycle_copy<Real>(set.cycleEntries, c0, nc, x->G, (Real*)set.cyclePsi,
                                         (Real*)set.psi, x->psiStride(), a0, true, x->stream));
    for (int h = set.blockLevel[b]; h < set.blockLevel[b + 1]; ++h) {
      if (timeLevels) HIP_TRY(hipEventRecord(set.evLevel[2 * h], x->stream));
      HIP_TRY(umt::launch_sweep_level<Real>(level_args<Real>(x, set, S, sigt, h, a0), false, x->stream));
      if (timeLevels) HIP_TRY(hipEventRecord(set.evLevel[2 * h + 1], x->stream));
    }
    // updateCycleList (SetSweep.F90:155)
    HIP_TRY(umt::launch_cycle_copy<Real>(set.cycleEntries, c0, nc, x->G, (Real*)set.cyclePsi,
                                         (Real*)set.psi, x->psiStride(), a0, false, x->stream));
    if (savePsi)  // Psi(:,:,Angle) = Psi1 for this block's angles (SweepUCBxyz.F90:314-318)
      HIP_TRY(hipMemcpyAsync((char*)savePsi + (size_t)a0 * x->psiStride() * sizeof(Real), set.psi,
                             (size_t)nA * x->psiStride() * sizeof(Real), hipMemcpyDeviceToDevice,
                             x->stream));
    HIP_TRY(umt::launch_phi_from_psi_acc<Real>((const Real*)set.psi, set.weight + a0, nA,
                                               x->psiStride(), (long)x->C * x->G, phiOut,
                                               accumulate || b > 0, x->stream));
  }
  return 0;
}

// One sweep's launches as a graph, replayed on later sweeps with the same
// source, cross section, groups per lane and output. The graph holds only
// kernel nodes (no memset or copy nodes).
template <typename Real>
int sweep_graph(umt_ctx* x, SweepSet& set, const void* S, const void* sigt, double* phiOut,
                bool accumulate, hipGraphExec_t* out) {
  for (auto& g : set.graphs)
    if (g.S == S && g.sigt == sigt && g.K == (set.groupsPerLane ? set.groupsPerLane : x->groupsPerLane) && g.phi == phiOut &&
        g.accumulate == accumulate) {
      *out = g.exec;
      return 0;
    }
  SweepSet::Graph g{S, sigt, set.groupsPerLane ? set.groupsPerLane : x->groupsPerLane, phiOut,
                    accumulate, nullptr, nullptr};
  HIP_TRY(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  if (int rc = launch_sweep<Real>(x, set, S, sigt, phiOut, accumulate, false)) {
    hipGraph_t dummy = nullptr;
    (void)hipStreamEndCapture(x->stream, &dummy);
    if (dummy) (void)hipGraphDestroy(dummy);
    return rc;
  }
  HIP_TRY(hipStreamEndCapture(x->stream, &g.graph));
  HIP_TRY(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  set.graphs.push_back(g);
  *out = g.exec;
  return 0;
}

// One sweep of every angle of `set` with source S and cross section sigt,
// then phiOut = (or +=) sum_a w_a psi.
template <typename Real>
int sweep_set_impl(umt_ctx* x, SweepSet& set, const void* S, const void* sigt, double* phiOut,
                   bool accumulate, bool timeLevels) {
  if (x->useGraphs && !timeLevels) {
    hipGraphExec_t exec = nullptr;
    if (int rc = sweep_graph<Real>(x, set, S, sigt, phiOut, accumulate, &exec)) return rc;
    HIP_TRY(hipGraphLaunch(exec, x->stream));
    return 0;
  }
  return launch_sweep<Real>(x, set, S, sigt, phiOut, accumulate, timeLevels);
}

int sweep_set(umt_ctx* x, SweepSet& set, const void* S, const void* sigt, double* phiOut,
              bool accumulate, bool timeLevels) {
  return x->precision == 32
             ? sweep_set_impl<float>(x, set, S, sigt, phiOut, accumulate, timeLevels)
             : sweep_set_impl<double>(x, set, S, sigt, phiOut, accumulate, timeLevels);
}

// The S_N sweep of the context: phi = sum_a w_a psi.
int sweep(umt_ctx* x, bool timeLevels) {
  return sweep_set(x, x->sn, x->S, x->sigt, x->phi, false, timeLevels);
}

// Upload a host FP64 array into a device array of the sweep precision.
int upload_real(umt_ctx* x, void* dst, const double* src, size_t n) {
  if (n == 0) return 0;
  if (x->precision == 64) {
    // Synchronous: callers pass temporaries (transposed copies).
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  } else {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  }
  return 0;
}

}  // namespace

extern "C" {

const char* umt_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Problem construction

}  // extern "C"

// Geometry, quadrature and both schedules of a problem whose mesh is built.
static int finish_problem(umt_problem* p, int npolar, int nazimuthal) {
  try {
    umt::compute_geometry(p->m, p->g);
    umt::product_quadrature(npolar, nazimuthal, p->q);
    std::string err;
    int rc = umt::build_schedule(p->m, p->g, p->q, p->s, err);
    if (rc) return fail(20 + rc, "schedule: %s", err.c_str());
    // The S2 directions lie in planar diagonal faces; on warped meshes their
    // schedule can have cycles, and then the problem has no GTA.
    umt::level_symmetric_s2(p->q2);
    if (umt::build_schedule(p->m, p->g, p->q2, p->s2, err, kGtaParallelTol)) {
      p->gtaError = "grey-acceleration schedule: " + err;
      p->s2 = umt::Schedule();
    } else if (p->s2.totalCycles() > 0) {
      // The grey solve is a Krylov method on a fixed linear operator; lagged
      // cycle fluxes would make the operator depend on the previous sweep.
      p->gtaError = "grey-acceleration schedule: the S2 directions have dependency cycles";
      p->s2 = umt::Schedule();
    }
  } catch (const std::exception& e) {
    return fail(2, "problem construction: %s", e.what());
  }
  return 0;
}

extern "C" {

int umt_problem_from_hex_mesh(int npoints, const double* coords, int nzones, const int* connectivity,
                              int npolar, int nazimuthal, int kdtree, umt_problem** out) {
  if (!out || !coords || !connectivity) return fail(1, "null argument");
  *out = nullptr;
  if (npoints < 8 || nzones < 1) return fail(1, "bad sizes");
  for (long i = 0; i < 3L * npoints; ++i)
    if (!std::isfinite(coords[i])) return fail(17, "coordinate %ld is not finite", i + 1);
  std::vector<std::array<int, 8>> hexes((size_t)nzones);
  for (long z = 0; z < nzones; ++z)
    for (int v = 0; v < 8; ++v) {
      const int id = connectivity[8 * z + v];
      if (id < 0 || id >= npoints) return fail(18, "zone %ld node %d: id %d out of range", z + 1, v + 1, id);
      hexes[z][v] = id;
    }
  umt_problem* p = new umt_problem();
  try {
    umt::build_hex_mesh(npoints, std::vector<double>(coords, coords + 3L * npoints), hexes,
                        kdtree != 0, p->m);
  } catch (const std::exception& e) {
    delete p;
    return fail(19, "mesh: %s", e.what());
  }
  if (int rc = finish_problem(p, npolar, nazimuthal)) {
    delete p;
    return rc;
  }
  *out = p;
  return 0;
}

int umt_problem_tiled(int nx, int ny, int nz, int npolar, int nazimuthal, int kdtree,
                      umt_problem** out) {
  if (!out) return fail(1, "null output pointer");
  *out = nullptr;
  umt_problem* p = new umt_problem();
  try {
    umt::build_tiled_mesh(nx, ny, nz, kdtree != 0, p->m);
  } catch (const std::exception& e) {
    delete p;
    return fail(2, "problem construction: %s", e.what());
  }
  if (int rc = finish_problem(p, npolar, nazimuthal)) {
    delete p;
    return rc;
  }
  *out = p;
  return 0;
}

void umt_problem_destroy(umt_problem* p) { delete p; }

}  // extern "C"

namespace umt {
const Mesh& problem_mesh(const umt_problem* p) { return p->m; }
const Geometry& problem_geometry(const umt_problem* p) { return p->g; }
const Quadrature& problem_quadrature(const umt_problem* p) { return p->q; }
const Schedule& problem_schedule(const umt_problem* p) { return p->s; }
}  // namespace umt

extern "C" {

int umt_problem_dims(const umt_problem* p, int64_t* d) {
  if (!p || !d) return fail(1, "null argument");
  d[0] = p->m.nzones;
  d[1] = p->m.ncorner;
  d[2] = p->m.nbelem;
  d[3] = p->q.nangles;
  d[4] = p->s.maxPlanes;
  d[5] = p->m.npoints;
  d[6] = (int64_t)p->s.zonesInPlane.size();
  d[7] = umt::kMaxCF;
  return 0;
}

static void copy_out(const std::vector<int>& v, int* dst) {
  if (dst) std::copy(v.begin(), v.end(), dst);
}
static void copy_out(const std::vector<double>& v, double* dst) {
  if (dst) std::copy(v.begin(), v.end(), dst);
}

int umt_problem_mesh(const umt_problem* p, int* numCorner, int* cOffSet, int* nCFaces, int* cFP,
                     int* cEZ, int* CToFace, int* zoneOpp, int* faceOpp, double* px) {
  if (!p) return fail(1, "null problem");
  copy_out(p->m.numCorner, numCorner);
  copy_out(p->m.cOffSet, cOffSet);
  copy_out(p->m.nCFaces, nCFaces);
  copy_out(p->m.cFP, cFP);
  copy_out(p->m.cEZ, cEZ);
  copy_out(p->m.CToFace, CToFace);
  copy_out(p->m.zoneOpp, zoneOpp);
  copy_out(p->m.faceOpp, faceOpp);
  copy_out(p->m.px, px);
  return 0;
}

int umt_problem_geometry(const umt_problem* p, double* A_fp, double* A_ez, double* Volume,
                         double* VolumeZone) {
  if (!p) return fail(1, "null problem");
  copy_out(p->g.A_fp, A_fp);
  copy_out(p->g.A_ez, A_ez);
  copy_out(p->g.Volume, Volume);
  copy_out(p->g.VolumeZone, VolumeZone);
  return 0;
}

int umt_problem_quadrature(const umt_problem* p, double* omega, double* weight) {
  if (!p) return fail(1, "null problem");
  copy_out(p->q.omega, omega);
  copy_out(p->q.weight, weight);
  return 0;
}

int umt_problem_schedule(const umt_problem* p, int* nHyperPlanes, int* zonesInPlane, int* nextZ,
                         int* nextC) {
  if (!p) return fail(1, "null problem");
  copy_out(p->s.nHyperPlanes, nHyperPlanes);
  copy_out(p->s.zonesInPlane, zonesInPlane);
  copy_out(p->s.nextZ, nextZ);
  copy_out(p->s.nextC, nextC);
  return 0;
}

int umt_problem_cycles(const umt_problem* p, int* totalCycles, int* numCycles, int* cycleList) {
  if (!p) return fail(1, "null problem");
  if (totalCycles) *totalCycles = (int)p->s.totalCycles();
  copy_out(p->s.numCycles, numCycles);
  copy_out(p->s.cycleList, cycleList);
  return 0;
}

int umt_problem_gta(const umt_problem* p, int* nangles, double* omega, double* weight,
                    int* nHyperPlanes, int* zonesInPlane, int* nextZ) {
  if (!p) return fail(1, "null problem");
  if (!p->gtaError.empty()) return fail(47, "%s", p->gtaError.c_str());
  if (nangles) *nangles = p->q2.nangles;
  copy_out(p->q2.omega, omega);
  copy_out(p->q2.weight, weight);
  copy_out(p->s2.nHyperPlanes, nHyperPlanes);
  copy_out(p->s2.zonesInPlane, zonesInPlane);
  copy_out(p->s2.nextZ, nextZ);
  return 0;
}

int umt_draw_cross_sections(uint64_t seed, int nzones, int groups, double sigt_lo, double sigt_hi,
                            double c_lo, double c_hi, double q_lo, double q_hi, double* sigt,
                            double* sigs, double* q) {
  if (nzones < 0 || groups < 1) return fail(1, "bad sizes");
  if (!(sigt_lo > 0.0) || !(sigt_hi >= sigt_lo)) return fail(1, "need 0 < sigt_lo <= sigt_hi");
  const double l0 = std::log(sigt_lo), l1 = std::log(sigt_hi);
  for (int z = 0; z < nzones; ++z)
    for (int g = 0; g < groups; ++g) {
      const size_t i = (size_t)z * groups + g;
      const double st = std::exp(l0 + umt::draw_uniform(seed, z, g, 0) * (l1 - l0));
      const double c = c_lo + umt::draw_uniform(seed, z, g, 1) * (c_hi - c_lo);
      if (sigt) sigt[i] = st;
      if (sigs) sigs[i] = c * st;
      if (q) q[i] = q_lo + umt::draw_uniform(seed, z, g, 2) * (q_hi - q_lo);
    }
  return 0;
}

// ---------------------------------------------------------------------------
// Device-resident context

static void drop_graphs(SweepSet& set) {
  for (auto& g : set.graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  set.graphs.clear();
}

static void free_set(SweepSet& set) {
  drop_graphs(set);
  dfree(set.omega);
  dfree(set.weight);
  dfree(set.items);
  dfree(set.psi);
  dfree(set.cycleEntries);
  dfree(set.cyclePsi);
  for (auto e : set.evLevel) (void)hipEventDestroy(e);
  set.evLevel.clear();
}

static void free_gta(GtaState& g) {
  dfree(g.sigt2); dfree(g.eps); dfree(g.S2); dfree(g.F); dfree(g.r); dfree(g.d); dfree(g.Ad);
  dfree(g.As); dfree(g.P1); dfree(g.gsrc); dfree(g.phiHalf); dfree(g.pzOld); dfree(g.partial);
  dfree(g.dots); dfree(g.coef); dfree(g.err); dfree(g.lu); dfree(g.localItems);
  g.luValid = false;
  g.enabled = false;
}

static void free_pending(LearnState& l) {
  for (auto& p : l.pending) {
    dfree(p.X);
    dfree(p.ph);
    dfree(p.d);
  }
  l.pending.clear();
}

static void free_learn(LearnState& l) {
  free_pending(l);
  dfree(l.zoneNbr); dfree(l.sigt64); dfree(l.pp); dfree(l.d); dfree(l.dp); dfree(l.rel);
  dfree(l.ratio); dfree(l.ta); dfree(l.tb); dfree(l.tpart); dfree(l.maxRel); dfree(l.scratchBits);
  dfree(l.X); dfree(l.r);
}

void umt_ctx_destroy(umt_ctx* x) {
  if (!x) return;
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  // Captured sweeps go first, then the buffers their kernel nodes name.
  drop_graphs(x->sn);
  drop_graphs(x->s2);
  free_learn(x->learn);
  dfree(x->psiPrev);
  dfree(x->psiInc);
  dfree(x->Afp); dfree(x->Aez); dfree(x->vol); dfree(x->cFP); dfree(x->vol64); dfree(x->volZ);
  dfree(x->S); dfree(x->sigt); dfree(x->sigs); dfree(x->q); dfree(x->phi); dfree(x->phiPrev);
  dfree(x->phibar); dfree(x->phibarOld); dfree(x->bits);
  free_set(x->sn);
  free_set(x->s2);
  free_gta(x->gta);
  if (x->ev0) (void)hipEventDestroy(x->ev0);
  if (x->ev1) (void)hipEventDestroy(x->ev1);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
}

// Hyperplane zone batches of one quadrature: level h holds plane h of every
// angle, angles in order, zones in each plane's nextZ order (block_set
// later reorders each level's items for the device; the order inside a level
// changes no result).
static int batch_levels(int Z, int A, const int* nHyperPlanes, const int* zonesInPlane,
                        const int* nextZ, std::vector<int2>& items, std::vector<int>& levelOffset,
                        bool* cyclicZones = nullptr) {
  if (cyclicZones) *cyclicZones = false;
  std::vector<int> planeStart(A + 1, 0);
  int maxPlanes = 0;
  for (int a = 0; a < A; ++a) {
    if (nHyperPlanes[a] < 1) return fail(16, "angle %d has no hyperplanes", a + 1);
    planeStart[a + 1] = planeStart[a] + nHyperPlanes[a];
    maxPlanes = std::max(maxPlanes, nHyperPlanes[a]);
  }
  std::vector<std::vector<int2>> lv(maxPlanes);
  for (int a = 0; a < A; ++a) {
    long done = 0;
    for (int h = 0; h < nHyperPlanes[a]; ++h) {
      const int n = zonesInPlane[planeStart[a] + h];
      if (n < 0 || done + n > Z) return fail(16, "angle %d: hyperplane sizes exceed the mesh", a + 1);
      for (int i = 0; i < n; ++i) {
        const int z = nextZ[(long)a * Z + done + i];
        if (z == 0 || z > Z || z < -Z) return fail(15, "angle %d: nextZ entry %d out of range", a + 1, z);
        // A cyclic zone (nextZ < 0, snac/snnext.F90:164-168) goes to the kernel as ~zone.
        lv[h].push_back(make_int2(z > 0 ? z - 1 : ~(-z - 1), a));
        if (z < 0 && cyclicZones) *cyclicZones = true;
      }
      done += n;
    }
    if (done != Z) return fail(16, "angle %d: hyperplanes hold %ld zones, mesh has %d", a + 1, done, Z);
  }
  items.clear();
  items.reserve((size_t)Z * A);
  levelOffset.assign(1, 0);
  for (int h = 0; h < maxPlanes; ++h) {
    items.insert(items.end(), lv[h].begin(), lv[h].end());
    levelOffset.push_back((int)items.size());
  }
  return 0;
}

// Splits the all-angle levels into angle blocks of n angles and (re)builds
// the device items, the psi buffer and the level events.
static int block_set(umt_ctx* x, SweepSet& set, int n) {
  n = std::max(1, std::min(n, set.A));
  const int nBlocks = (set.A + n - 1) / n;
  std::vector<int2> items;
  items.reserve(set.allItems.size());
  std::vector<int> lo(1, 0), bl(1, 0);
  const int L = (int)set.allLevelOffset.size() - 1;
  for (int b = 0; b < nBlocks; ++b) {
    const int a0 = b * n, a1 = std::min(set.A, a0 + n);
    for (int h = 0; h < L; ++h) {
      const size_t before = items.size();
      for (int i = set.allLevelOffset[h]; i < set.allLevelOffset[h + 1]; ++i) {
        const int2 it = set.allItems[i];
        if (it.y >= a0 && it.y < a1) items.push_back(it);
      }
      if (items.size() > before) lo.push_back((int)items.size());
    }
    bl.push_back((int)lo.size() - 1);
  }
  // A level's items in zone order (then angle): the directions that reach a
  // zone at the same level (on hex tiles, classes of directions share their
  // hyperplanes) sit side by side, so their reads of the zone's source,
  // cross section and corner geometry hit the same cache lines. Items of a
  // level are independent, so the order changes no result.
  long zoneVisits = 0;
  for (size_t h = 0; h + 1 < lo.size(); ++h) {
    std::sort(items.begin() + lo[h], items.begin() + lo[h + 1], [](int2 p, int2 q) {
      const int zp = p.x < 0 ? ~p.x : p.x, zq = q.x < 0 ? ~q.x : q.x;
      return zp != zq ? zp < zq : p.y < q.y;
    });
    for (int i = lo[h]; i < lo[h + 1]; ++i) {
      const int z = items[i].x < 0 ? ~items[i].x : items[i].x;
      if (i == lo[h] || z != (items[i - 1].x < 0 ? ~items[i - 1].x : items[i - 1].x)) ++zoneVisits;
    }
  }
  // XCD-aware order: workgroup w of a level launch runs on XCD w mod 8, so
  // slot w takes the next workgroup's worth of items from zone range w mod 8
  // (eight equal ranges of zone ids), or from the fullest range once its own
  // is used up. A zone and its upwind neighbours then mostly meet in the same
  // XCD's L2 from level to level.
  {
    constexpr int kXcd = 8;
    // Items per 256-lane workgroup of the layout the autotuner usually
    // picks: the zone layout (one lane per group) where G allows it, else
    // the widest groups-per-lane layout; another layout only mixes ranges.
    const int K = x->G % 4 == 0 ? 4 : (x->G % 2 == 0 ? 2 : 1);
    const int perWg = umt::zone_lanes_for(x->G) ? std::max(1, 256 / x->G)
                                                : std::max(1, 32 / std::max(1, x->G / K));
    std::vector<int2> out;
    for (size_t h = 0; h + 1 < lo.size(); ++h) {
      std::array<int, kXcd + 1> start;
      start[0] = lo[h];
      for (int r = 1; r < kXcd; ++r) {
        const int zcut = (int)((long)x->Z * r / kXcd);
        start[r] = (int)(std::lower_bound(items.begin() + start[r - 1], items.begin() + lo[h + 1], zcut,
                                          [](int2 it, int zc) { return (it.x < 0 ? ~it.x : it.x) < zc; }) -
                         items.begin());
      }
      start[kXcd] = lo[h + 1];
      std::array<int, kXcd> next;
      for (int r = 0; r < kXcd; ++r) next[r] = start[r];
      out.clear();
      for (long slot = 0; (int)out.size() < lo[h + 1] - lo[h]; ++slot) {
        int r = (int)(slot % kXcd);
        if (next[r] == start[r + 1]) {
          int best = -1;
          for (int t = 0; t < kXcd; ++t)
            if (next[t] < start[t + 1] && (best < 0 || start[t + 1] - next[t] > start[best + 1] - next[best]))
              best = t;
          r = best;
        }
        const int take = std::min(perWg, start[r + 1] - next[r]);
        out.insert(out.end(), items.begin() + next[r], items.begin() + next[r] + take);
        next[r] += take;
      }
      std::copy(out.begin(), out.end(), items.begin() + lo[h]);
    }
  }
  if (x->stream) HIP_TRY(hipStreamSynchronize(x->stream));
  drop_graphs(set);
  for (auto e : set.evLevel) (void)hipEventDestroy(e);
  set.evLevel.clear();
  dfree(set.items);
  dfree(set.psi);
  const size_t nPsi = (size_t)x->psiStride() * n;
  int rc;
  if ((rc = dalloc(&set.items, items.size())) || (rc = dalloc((char**)&set.psi, nPsi * x->rsize())))
    return rc;
  HIP_TRY(hipMemcpyAsync(set.items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, x->stream));
  if (x->precision == 32)
    HIP_TRY(umt::launch_fill_zero<float>((float*)set.psi, (long)nPsi, x->stream));
  else
    HIP_TRY(umt::launch_fill_zero<double>((double*)set.psi, (long)nPsi, x->stream));
  set.anglesPerBlock = n;
  set.zoneVisits = zoneVisits;
  set.levelOffset = lo;
  set.blockLevel = bl;
  set.evLevel.resize(2 * (size_t)set.levels());
  for (auto& e : set.evLevel) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

// Levels of angles with cycle lists. Zones of one plane can then be face
// neighbours (a link broken by the cycle breaker whose two zones were
// released together, or a face with no net flow), and the serial sweep order
// inside the plane decides whether a zone reads its neighbour's new or
// lagged psi. Such a plane is split into sub-levels in nextZ order: a zone
// goes one sub-level after the latest earlier zone of its plane that it
// touches, so every read sees what the serial sweep sees. Level L of the
// result is sub-level L of every angle. Idempotent.
static void split_cyclic_planes(int Z, int C, const int* cFP, int A,
                                const std::vector<int>& numCycles, std::vector<int2>& items,
                                std::vector<int>& levelOffset) {
  std::vector<std::vector<int2>> out;
  std::vector<int> next(A, 0);  // next output level of each angle
  std::vector<int> mark(Z, -1), sub;
  auto put = [&](int level, int2 it) {
    if ((int)out.size() <= level) out.resize(level + 1);
    out[level].push_back(it);
  };
  const int L = (int)levelOffset.size() - 1;
  for (int h = 0; h < L; ++h) {
    int i = levelOffset[h];
    while (i < levelOffset[h + 1]) {
      const int a = items[i].y;
      int j = i;
      while (j < levelOffset[h + 1] && items[j].y == a) ++j;
      if (numCycles[a] == 0) {
        for (int k = i; k < j; ++k) put(next[a], items[k]);
        next[a] += 1;
      } else {
        sub.assign(j - i, 0);
        int maxSub = 0;
        for (int k = i; k < j; ++k) {
          const int z = items[k].x < 0 ? ~items[k].x : items[k].x;
          int sl = 0;
          for (int c = 8 * z; c < 8 * z + 8; ++c)
            for (int f = 0; f < 3; ++f) {
              const int nb = cFP[3 * (size_t)c + f] - 1;
              if (nb >= C) continue;
              const int zn = nb / 8;
              if (zn != z && mark[zn] >= 0) sl = std::max(sl, sub[mark[zn]] + 1);
            }
          sub[k - i] = sl;
          mark[z] = k - i;
          maxSub = std::max(maxSub, sl);
        }
        for (int s = 0; s <= maxSub; ++s)
          for (int k = i; k < j; ++k)
            if (sub[k - i] == s) put(next[a] + s, items[k]);
        for (int k = i; k < j; ++k) mark[items[k].x < 0 ? ~items[k].x : items[k].x] = -1;
        next[a] += maxSub + 1;
      }
      i = j;
    }
  }
  items.clear();
  levelOffset.assign(1, 0);
  for (auto& lv : out) {
    items.insert(items.end(), lv.begin(), lv.end());
    levelOffset.push_back((int)items.size());
  }
}

static int upload_set(umt_ctx* x, SweepSet& set, int A, const double* omega, const double* weight,
                      const int* nHyperPlanes, const int* zonesInPlane, const int* nextZ) {
  if (int rc = batch_levels(x->Z, A, nHyperPlanes, zonesInPlane, nextZ, set.allItems,
                            set.allLevelOffset, &set.cyclicZones))
    return rc;
  set.A = A;
  int rc;
  if ((rc = dalloc(&set.omega, 3 * (size_t)A)) || (rc = dalloc(&set.weight, (size_t)A))) return rc;
  HIP_TRY(hipMemcpyAsync(set.omega, omega, 3 * (size_t)A * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(set.weight, weight, (size_t)A * sizeof(double), hipMemcpyHostToDevice, x->stream));
  return block_set(x, set, A);
}

static int ctx_create_impl(int Z, int C, int nb, int G, int A, int precision, const int* numCorner,
                           const int* cOffSet, const int* nCFaces, const int* cFP, const int* cEZ,
                           const double* A_fp, const double* A_ez, const double* Volume,
                           const double* VolumeZone, const double* omega, const double* weight,
                           const int* nHyperPlanes, const int* zonesInPlane, const int* nextZ,
                           umt_ctx* x) {
  if (Z < 1 || C < 1 || nb < 0 || G < 1 || A < 1) return fail(1, "bad sizes");
  if (precision != 32 && precision != 64) return fail(1, "precision must be 32 or 64");
  if (int rc = check_hex_topology(Z, C, numCorner, cOffSet, nCFaces, cEZ)) return rc;
  for (long i = 0; i < 3L * C; ++i)
    if (cFP[i] < 1 || cFP[i] > C + nb) return fail(14, "cFP entry %ld out of range", i + 1);
  {  // validate the schedule before touching the device
    std::vector<int2> items;
    std::vector<int> lo;
    if (int rc = batch_levels(Z, A, nHyperPlanes, zonesInPlane, nextZ, items, lo)) return rc;
  }
  // n_in of the algorithmic-bytes model, counted exactly.
  {
    long long cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
    for (int a = 0; a < A; ++a) {
      const double* o = omega + 3L * a;
      for (long i = 0; i < 3L * C; ++i) {
        const double* v = A_fp + 3 * i;
        if (o[0] * v[0] + o[1] * v[1] + o[2] * v[2] < 0.0 && cFP[i] <= C) ++cnt;
      }
    }
    x->nIn = (double)cnt / ((double)C * A);
  }

  x->Z = Z; x->C = C; x->nb = nb; x->G = G; x->precision = precision;
  HIP_TRY(hipStreamCreate(&x->stream));
  HIP_TRY(hipEventCreate(&x->ev0));
  HIP_TRY(hipEventCreate(&x->ev1));

  const size_t rs = x->rsize();
  int rc = 0;
  if ((rc = dalloc((char**)&x->Afp, 9 * (size_t)C * rs)) || (rc = dalloc((char**)&x->Aez, 9 * (size_t)C * rs)) ||
      (rc = dalloc((char**)&x->vol, (size_t)C * rs)) || (rc = dalloc(&x->cFP, 3 * (size_t)C)) ||
      (rc = dalloc(&x->vol64, (size_t)C)) || (rc = dalloc(&x->volZ, (size_t)Z)) ||
      (rc = dalloc((char**)&x->S, (size_t)G * C * rs)) ||
      (rc = dalloc((char**)&x->sigt, (size_t)G * Z * rs)) || (rc = dalloc(&x->sigs, (size_t)G * Z)) ||
      (rc = dalloc(&x->q, (size_t)G * Z)) || (rc = dalloc(&x->phi, (size_t)G * C)) ||
      (rc = dalloc(&x->phiPrev, (size_t)G * C)) || (rc = dalloc(&x->phibar, (size_t)G * Z)) ||
      (rc = dalloc(&x->phibarOld, (size_t)G * Z)) || (rc = dalloc(&x->bits, 2)))
    return rc;

  if ((rc = upload_real(x, x->Afp, umt::zone_blocked(A_fp, 9, (size_t)C).data(), 9 * (size_t)C)) ||
      (rc = upload_real(x, x->Aez, umt::zone_blocked(A_ez, 9, (size_t)C).data(), 9 * (size_t)C)) ||
      (rc = upload_real(x, x->vol, Volume, (size_t)C)))
    return rc;
  const std::vector<int> cFPdev = umt::zone_blocked(cFP, 3, (size_t)C);
  HIP_TRY(hipMemcpyAsync(x->cFP, cFPdev.data(), 3 * (size_t)C * sizeof(int), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  x->cFPHost.assign(cFP, cFP + 3 * (size_t)C);
  HIP_TRY(hipMemcpyAsync(x->vol64, Volume, (size_t)C * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(x->volZ, VolumeZone, (size_t)Z * sizeof(double), hipMemcpyHostToDevice, x->stream));
  if (precision == 32) {
    HIP_TRY(umt::launch_fill_zero<float>((float*)x->S, (long)G * C, x->stream));
    HIP_TRY(umt::launch_fill_zero<float>((float*)x->sigt, (long)G * Z, x->stream));
  } else {
    HIP_TRY(umt::launch_fill_zero<double>((double*)x->S, (long)G * C, x->stream));
    HIP_TRY(umt::launch_fill_zero<double>((double*)x->sigt, (long)G * Z, x->stream));
  }
  HIP_TRY(umt::launch_fill_zero<double>(x->phi, (long)G * C, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>(x->phibarOld, (long)G * Z, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>(x->sigs, (long)G * Z, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>(x->q, (long)G * Z, x->stream));
  if ((rc = upload_set(x, x->sn, A, omega, weight, nHyperPlanes, zonesInPlane, nextZ))) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_create(int nzones, int ncorner, int nbelem, int groups, int nangles, int precision,
                   const int* numCorner, const int* cOffSet, const int* nCFaces, const int* cFP,
                   const int* cEZ, const double* A_fp, const double* A_ez, const double* Volume,
                   const double* VolumeZone, const double* omega, const double* weight,
                   const int* nHyperPlanes, const int* zonesInPlane, const int* nextZ,
                   umt_ctx** out) {
  if (!out) return fail(1, "null output pointer");
  *out = nullptr;
  umt_ctx* x = new umt_ctx();
  int rc = ctx_create_impl(nzones, ncorner, nbelem, groups, nangles, precision, numCorner, cOffSet,
                           nCFaces, cFP, cEZ, A_fp, A_ez, Volume, VolumeZone, omega, weight,
                           nHyperPlanes, zonesInPlane, nextZ, x);
  if (rc) {
    std::string keep = g_err;
    umt_ctx_destroy(x);
    g_err = keep;
    return rc;
  }
  *out = x;
  return 0;
}

int umt_ctx_create_from_problem(const umt_problem* p, int groups, int precision, umt_ctx** out) {
  if (!p) return fail(1, "null problem");
  int rc = umt_ctx_create(p->m.nzones, p->m.ncorner, p->m.nbelem, groups, p->q.nangles, precision,
                          p->m.numCorner.data(), p->m.cOffSet.data(), p->m.nCFaces.data(),
                          p->m.cFP.data(), p->m.cEZ.data(), p->g.A_fp.data(), p->g.A_ez.data(),
                          p->g.Volume.data(), p->g.VolumeZone.data(), p->q.omega.data(),
                          p->q.weight.data(), p->s.nHyperPlanes.data(), p->s.zonesInPlane.data(),
                          p->s.nextZ.data(), out);
  if (rc) return rc;
  if (p->s.totalCycles() > 0)
    rc = umt_ctx_set_cycles(*out, p->s.numCycles.data(), p->s.cycleList.data());
  if (!rc && p->gtaError.empty())
    rc = umt_ctx_enable_gta(*out, p->q2.nangles, p->q2.omega.data(), p->q2.weight.data(),
                            p->s2.nHyperPlanes.data(), p->s2.zonesInPlane.data(),
                            p->s2.nextZ.data());
  if (!rc) rc = umt_ctx_set_zone_neighbors(*out, p->m.zoneOpp.data());
  if (rc) {
    std::string keep = g_err;
    umt_ctx_destroy(*out);
    *out = nullptr;
    g_err = keep;
  }
  return rc;
}

int umt_ctx_dims(const umt_ctx* x, int64_t* d) {
  if (!x || !d) return fail(1, "null argument");
  d[0] = x->Z; d[1] = x->C; d[2] = x->nb; d[3] = x->G; d[4] = x->sn.A;
  d[5] = x->sn.levels();
  d[6] = x->sn.levelOffset.back();
  return 0;
}

// The sweep's total cross section: sigma_t + tau (control/setTotalOpacity.F90).
static int upload_sigt(umt_ctx* x) {
  std::vector<double> st(x->sigtHost);
  for (double& v : st) v += x->tau;
  int rc = upload_real(x, x->sigt, st.data(), st.size());
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_set_sigt(umt_ctx* x, const double* sigt) {
  if (!x || !sigt) return fail(1, "null argument");
  for (long i = 0; i < (long)x->G * x->Z; ++i)
    if (!(sigt[i] > 0.0) || !std::isfinite(sigt[i])) return fail(30, "Sigt entry %ld is not positive and finite", i + 1);
  x->sigtHost.assign(sigt, sigt + (size_t)x->G * x->Z);
  if (int rc = upload_sigt(x)) return rc;
  x->haveSigt = true;
  return 0;
}

int umt_ctx_set_tau(umt_ctx* x, double tau) {
  if (!x) return fail(1, "null context");
  if (!(tau >= 0.0) || !std::isfinite(tau)) return fail(48, "tau must be finite and >= 0");
  HIP_TRY(hipStreamSynchronize(x->stream));
  if (tau > 0.0 && !x->psiPrev) {
    const size_t n = (size_t)x->psiStride() * x->sn.A;
    if (int rc = dalloc((char**)&x->psiPrev, n * x->rsize())) return rc;
    if (x->precision == 32) HIP_TRY(umt::launch_fill_zero<float>((float*)x->psiPrev, (long)n, x->stream));
    else HIP_TRY(umt::launch_fill_zero<double>((double*)x->psiPrev, (long)n, x->stream));
  }
  x->tau = tau;
  drop_graphs(x->sn);  // the level launches carry tau and the Psi pointer
  if (x->haveSigt)
    if (int rc = upload_sigt(x)) return rc;
  return 0;
}

int umt_ctx_set_incident_flux(umt_ctx* x, const double* PsiB) {
  if (!x) return fail(1, "null context");
  HIP_TRY(hipStreamSynchronize(x->stream));
  drop_graphs(x->sn);  // captured sweeps hold no boundary copies
  if (!PsiB) {
    dfree(x->psiInc);
    x->psiInc = nullptr;
    return 0;
  }
  const size_t n = (size_t)x->nb * x->G * x->sn.A;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(PsiB[i])) return fail(56, "PsiB entry %zu is not finite", i + 1);
  if (n == 0) return 0;
  if (!x->psiInc)
    if (int rc = dalloc((char**)&x->psiInc, n * x->rsize())) return rc;
  return upload_real(x, x->psiInc, PsiB, n);
}

int umt_ctx_get_boundary_flux(const umt_ctx* x, double* PsiB) {
  if (!x || !PsiB) return fail(1, "null argument");
  if (x->sn.blocks() != 1)
    return fail(45, "boundary fluxes of every angle are held only with one angle block");
  const size_t nbG = (size_t)x->nb * x->G, CG = (size_t)x->C * x->G;
  HIP_TRY(hipStreamSynchronize(x->stream));
  for (int a = 0; a < x->sn.A; ++a) {
    const size_t off = (size_t)a * x->psiStride() + CG;
    if (x->precision == 64) {
      HIP_TRY(hipMemcpy(PsiB + a * nbG, (const double*)x->sn.psi + off, nbG * sizeof(double),
                        hipMemcpyDeviceToHost));
    } else {
      std::vector<float> f(nbG);
      HIP_TRY(hipMemcpy(f.data(), (const float*)x->sn.psi + off, nbG * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < nbG; ++i) PsiB[a * nbG + i] = f[i];
    }
  }
  return 0;
}

int umt_ctx_set_psi_prev(umt_ctx* x, const double* psi) {
  if (!x || !psi) return fail(1, "null argument");
  if (!x->psiPrev) return fail(49, "umt_ctx_set_tau with tau > 0 first");
  const size_t GC = (size_t)x->G * x->C, stride = (size_t)x->psiStride();
  std::vector<double> h(stride * x->sn.A, 0.0);
  for (int a = 0; a < x->sn.A; ++a) std::memcpy(&h[a * stride], psi + a * GC, GC * sizeof(double));
  int rc = upload_real(x, x->psiPrev, h.data(), h.size());
  if (rc) return rc;
  // initCyclePsi (control/constructDynMemory.F90:75-114): the lagged values
  // of a new time step start from the previous cycle's psi.
  const SweepSet& sn = x->sn;
  if (x->precision == 32)
    HIP_TRY(umt::launch_cycle_copy<float>(sn.cycleEntries, 0, sn.totalCycles(), x->G,
                                          (float*)sn.cyclePsi, (float*)x->psiPrev,
                                          x->psiStride(), 0, false, x->stream));
  else
    HIP_TRY(umt::launch_cycle_copy<double>(sn.cycleEntries, 0, sn.totalCycles(), x->G,
                                           (double*)sn.cyclePsi, (double*)x->psiPrev,
                                           x->psiStride(), 0, false, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_get_psi_prev(const umt_ctx* x, double* psi) {
  if (!x || !psi) return fail(1, "null argument");
  if (!x->psiPrev) return fail(49, "umt_ctx_set_tau with tau > 0 first");
  const size_t GC = (size_t)x->G * x->C, stride = (size_t)x->psiStride();
  const size_t n = stride * x->sn.A;
  std::vector<double> h(n);
  if (x->precision == 64) {
    HIP_TRY(hipMemcpyAsync(h.data(), x->psiPrev, n * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  } else {
    std::vector<float> f(n);
    HIP_TRY(hipMemcpyAsync(f.data(), x->psiPrev, n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    for (size_t i = 0; i < n; ++i) h[i] = f[i];
  }
  for (int a = 0; a < x->sn.A; ++a) std::memcpy(psi + a * GC, &h[a * stride], GC * sizeof(double));
  return 0;
}

int umt_ctx_set_source(umt_ctx* x, const double* S) {
  if (!x || !S) return fail(1, "null argument");
  for (long i = 0; i < (long)x->G * x->C; ++i)
    if (!std::isfinite(S[i])) return fail(31, "STotal entry %ld is not finite", i + 1);
  int rc = upload_real(x, x->S, S, (size_t)x->G * x->C);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_set_scattering(umt_ctx* x, const double* sigs, const double* q) {
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

int umt_ctx_sweep(umt_ctx* x) {
  if (!x) return fail(1, "null context");
  if (int rc = sweep(x, false)) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_sweep_n(umt_ctx* x, int n) {
  if (!x || n < 0) return fail(1, "bad argument");
  for (int i = 0; i < n; ++i)
    if (int rc = sweep(x, false)) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_set_groups_per_lane(umt_ctx* x, int K) {
  if (!x) return fail(1, "null context");
  if (!umt::sweep_layout_supported(x->G, K))
    return fail(35, "sweep layout %d not supported for %d groups (groups per lane 1, 2, 4 or 8 "
                    "dividing G, or -1: zone layout, G a multiple of 16)", K, x->G);
  x->groupsPerLane = K;
  return 0;
}

int umt_ctx_set_graphs(umt_ctx* x, int on) {
  if (!x) return fail(1, "null context");
  x->useGraphs = on != 0;
  return 0;
}

int umt_ctx_set_angles_per_block(umt_ctx* x, int n) {
  if (!x) return fail(1, "null context");
  if (n < 1) return fail(1, "angles per block must be at least 1");
  return block_set(x, x->sn, n);
}

int umt_ctx_get_angles_per_block(const umt_ctx* x, int* n, int* blocks) {
  if (!x) return fail(1, "null context");
  if (n) *n = x->sn.anglesPerBlock;
  if (blocks) *blocks = x->sn.blocks();
  return 0;
}

int umt_ctx_copy_phi(umt_ctx* dst, const umt_ctx* src) {
  if (!dst || !src) return fail(1, "null context");
  if (dst->G != src->G || dst->C != src->C) return fail(46, "contexts differ in groups or corners");
  HIP_TRY(hipStreamSynchronize(src->stream));
  HIP_TRY(hipMemcpyAsync(dst->phi, src->phi, (size_t)src->G * src->C * sizeof(double),
                         hipMemcpyDeviceToDevice, dst->stream));
  // The lagged cycle psi continues too, converted to the destination's precision.
  const long nc = (long)src->G * src->sn.totalCycles();
  if (nc > 0 && src->sn.cycleOffset == dst->sn.cycleOffset) {
    const void* a = src->sn.cyclePsi;
    void* b = dst->sn.cyclePsi;
    if (src->precision == 64 && dst->precision == 64)
      HIP_TRY((umt::launch_convert<double, double>((const double*)a, (double*)b, nc, dst->stream)));
    else if (src->precision == 64)
      HIP_TRY((umt::launch_convert<float, double>((const double*)a, (float*)b, nc, dst->stream)));
    else if (dst->precision == 64)
      HIP_TRY((umt::launch_convert<double, float>((const float*)a, (double*)b, nc, dst->stream)));
    else
      HIP_TRY((umt::launch_convert<float, float>((const float*)a, (float*)b, nc, dst->stream)));
  }
  HIP_TRY(hipStreamSynchronize(dst->stream));
  return 0;
}

int umt_ctx_set_cycles(umt_ctx* x, const int* numCycles, const int* cycleList) {
  if (!x || !numCycles) return fail(1, "null argument");
  SweepSet& set = x->sn;
  const int A = set.A;
  std::vector<int> offset(A + 1, 0);
  for (int a = 0; a < A; ++a) {
    if (numCycles[a] < 0 || numCycles[a] > x->C)
      return fail(50, "angle %d: numCycles = %d out of range", a + 1, numCycles[a]);
    offset[a + 1] = offset[a] + numCycles[a];
  }
  const long total = offset[A];
  if (total > 0 && !cycleList) return fail(1, "null cycle list");
  std::vector<int2> entries((size_t)total);
  for (int a = 0; a < A; ++a)
    for (int m = offset[a]; m < offset[a + 1]; ++m) {
      const int c = cycleList[m];
      if (c < 1 || c > x->C) return fail(50, "angle %d: cycle list corner %d out of range", a + 1, c);
      entries[m] = make_int2(c - 1, a);
    }
  HIP_TRY(hipStreamSynchronize(x->stream));
  dfree(set.cycleEntries);
  dfree(set.cyclePsi);
  set.cycleOffset.clear();
  if (total > 0) {
    int rc;
    if ((rc = dalloc(&set.cycleEntries, (size_t)total)) ||
        (rc = dalloc((char**)&set.cyclePsi, (size_t)total * x->G * x->rsize())))
      return rc;
    HIP_TRY(hipMemcpyAsync(set.cycleEntries, entries.data(), (size_t)total * sizeof(int2),
                           hipMemcpyHostToDevice, x->stream));
    // Lagged values start at zero, as SetData's cyclePsi does before
    // initCyclePsi loads a previous cycle's psi.
    if (x->precision == 32)
      HIP_TRY(umt::launch_fill_zero<float>((float*)set.cyclePsi, total * x->G, x->stream));
    else
      HIP_TRY(umt::launch_fill_zero<double>((double*)set.cyclePsi, total * x->G, x->stream));
    set.cycleOffset = offset;
  }
  std::vector<int> nc(numCycles, numCycles + A);
  split_cyclic_planes(x->Z, x->C, x->cFPHost.data(), A, nc, set.allItems, set.allLevelOffset);
  return block_set(x, set, set.anglesPerBlock);
}

int umt_ctx_get_groups_per_lane(const umt_ctx* x, int* K) {
  if (!x || !K) return fail(1, "null argument");
  *K = x->groupsPerLane;
  return 0;
}

int umt_ctx_autotune_layouts(umt_ctx* x, int reps, int* bestK, double* msPerLayout) {
  if (!x || reps < 1) return fail(1, "bad argument");
  static const int cand[5] = {1, 2, 4, 8, umt::kZoneLayout};
  double best = 1e300;
  int bk = x->groupsPerLane;
  for (int i = 0; i < 5; ++i) {
    if (msPerLayout) msPerLayout[i] = -1.0;
    if (!umt::sweep_layout_supported(x->G, cand[i])) continue;
    x->groupsPerLane = cand[i];
    if (int rc = sweep(x, false)) return rc;  // warm-up
    HIP_TRY(hipEventRecord(x->ev0, x->stream));
    for (int r = 0; r < reps; ++r)
      if (int rc = sweep(x, false)) return rc;
    HIP_TRY(hipEventRecord(x->ev1, x->stream));
    HIP_TRY(hipEventSynchronize(x->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, x->ev0, x->ev1));
    const double ms = t / reps;
    if (msPerLayout) msPerLayout[i] = ms;
    if (ms < best) {
      best = ms;
      bk = cand[i];
    }
  }
  x->groupsPerLane = bk;
  if (bestK) *bestK = bk;
  return 0;
}

int umt_ctx_autotune(umt_ctx* x, int reps, int* bestK, double* msPerK) {
  double ms[5];
  if (int rc = umt_ctx_autotune_layouts(x, reps, bestK, ms)) return rc;
  if (msPerK)
    for (int i = 0; i < 4; ++i) msPerK[i] = ms[i];
  return 0;
}

int umt_ctx_incoming_faces(const umt_ctx* x, double* n_in) {
  if (!x || !n_in) return fail(1, "null argument");
  *n_in = x->nIn;
  return 0;
}

int umt_device_memory_info(int64_t* freeBytes, int64_t* totalBytes) {
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  if (freeBytes) *freeBytes = (int64_t)f;
  if (totalBytes) *totalBytes = (int64_t)t;
  return 0;
}

int umt_ctx_level_stats(const umt_ctx* x, int64_t* stats) {
  if (!x || !stats) return fail(1, "null argument");
  stats[0] = x->sn.levels();
  stats[1] = x->sn.levelOffset.back();
  stats[2] = x->sn.zoneVisits;
  stats[3] = x->sn.blocks();
  return 0;
}

int umt_hbm_copy_probe(int64_t n, int reps, double* ms) {
  if (n < 1 || reps < 1 || !ms) return fail(1, "bad argument");
  double *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = dalloc(&a, (size_t)n);
  if (!rc) rc = dalloc(&b, (size_t)n);
  if (!rc) {
    HIP_TRY(hipStreamCreate(&s));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(umt::launch_fill_zero<double>(a, (long)n, s));
    HIP_TRY(umt::launch_copy_probe(a, b, (long)n, s));
    HIP_TRY(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) HIP_TRY(umt::launch_copy_probe(a, b, (long)n, s));
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  dfree(a);
  dfree(b);
  return rc;
}

int umt_ctx_get_phi(const umt_ctx* x, double* phi) {
  if (!x || !phi) return fail(1, "null argument");
  HIP_TRY(hipMemcpyAsync(phi, x->phi, (size_t)x->G * x->C * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_set_phi(umt_ctx* x, const double* phi) {
  if (!x || !phi) return fail(1, "null argument");
  HIP_TRY(hipMemcpyAsync(x->phi, phi, (size_t)x->G * x->C * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_get_psi(const umt_ctx* x, double* psi) {
  if (!x || !psi) return fail(1, "null argument");
  if (x->sn.blocks() != 1)
    return fail(45, "psi of every angle is held only with one angle block (umt_ctx_set_angles_per_block)");
  const size_t n = (size_t)x->psiStride() * x->sn.A;
  if (x->precision == 64) {
    HIP_TRY(hipMemcpyAsync(psi, x->sn.psi, n * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  } else {
    std::vector<float> tmp(n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), x->sn.psi, n * sizeof(float), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    for (size_t i = 0; i < n; ++i) psi[i] = tmp[i];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Grey-transport acceleration, one one-group GTA problem per group
// (rt/GTASolver.F90, rt/GreySweep.F90, snac/GTASweep.F90).

int umt_ctx_enable_gta(umt_ctx* x, int nangles, const double* omega, const double* weight,
                       const int* nHyperPlanes, const int* zonesInPlane, const int* nextZ) {
  if (!x || !omega || !weight || !nHyperPlanes || !zonesInPlane || !nextZ) return fail(1, "null argument");
  if (nangles < 1) return fail(1, "bad angle count");
  if (x->gta.enabled) {  // replaces the S2 set: no sweep may still read it
    HIP_TRY(hipStreamSynchronize(x->stream));
    free_set(x->s2);
    free_gta(x->gta);
  }
  if (int rc = upload_set(x, x->s2, nangles, omega, weight, nHyperPlanes, zonesInPlane, nextZ)) return rc;
  x->s2.parTol = x->precision == 32 ? kGtaParallelTolFP32 : kGtaParallelTol;
  GtaState& g = x->gta;
  const size_t GC = (size_t)x->G * x->C, GZ = (size_t)x->G * x->Z, rs = x->rsize();
  g.dotBlocks = 512;
  g.errBlocks = 512;
  int rc;
  if ((rc = dalloc((char**)&g.sigt2, 2 * GZ * rs)) || (rc = dalloc(&g.eps, 2 * GZ)) ||
      (rc = dalloc((char**)&g.S2, GC * rs)) || (rc = dalloc(&g.F, GC)) || (rc = dalloc(&g.r, GC)) ||
      (rc = dalloc(&g.d, GC)) || (rc = dalloc(&g.Ad, GC)) || (rc = dalloc(&g.As, GC)) ||
      (rc = dalloc(&g.P1, GC)) || (rc = dalloc(&g.gsrc, GC)) || (rc = dalloc(&g.phiHalf, GC)) ||
      (rc = dalloc(&g.pzOld, GZ)) ||
      (rc = dalloc(&g.partial, (size_t)std::max(2 * g.dotBlocks * x->G, 3 * g.errBlocks))) ||
      (rc = dalloc(&g.dots, 4 * (size_t)x->G)) || (rc = dalloc(&g.coef, 3 * (size_t)x->G)) ||
      (rc = dalloc(&g.err, 3)))
    return rc;
  g.enabled = true;
  return 0;
}

int umt_ctx_set_acceleration(umt_ctx* x, int mode) {
  if (!x) return fail(1, "null context");
  if (mode < 0 || mode > 3)
    return fail(36, "acceleration must be 0 (none), 1 (GTA), 2 (learned) or 3 (GTA + learned)");
  if ((mode == 1 || mode == 3) && !x->gta.enabled) return fail(37, "GTA needs umt_ctx_enable_gta first");
  if (mode >= 2 && !x->learn.net) return fail(38, "learned acceleration needs umt_ctx_set_learner first");
  if (mode >= 2 && !x->learn.zoneNbr) return fail(39, "learned acceleration needs umt_ctx_set_zone_neighbors first");
  x->acceleration = mode;
  return 0;
}

int umt_ctx_set_gta_solver(umt_ctx* x, int kernel) {
  if (!x) return fail(1, "null context");
  if (kernel != 1 && kernel != 2) return fail(55, "GTA kernel must be 1 (grey + stretched TSA) or 2 (transfer matrices)");
  x->gta.solver = kernel;
  x->gta.hardIterMax = kernel == 2;  // each kernel's default; umt_ctx_set_gta_options overrides
  return 0;
}

int umt_ctx_set_grey_controls(umt_ctx* x, double relTol, int maxIts) {
  if (!x) return fail(1, "null context");
  if (!(relTol > 0.0) || maxIts < 1)
    return fail(57, "grey controls need a relative tolerance > 0 and at least one iteration");
  x->gta.epsPoint = relTol;
  x->gta.maxSweeps = maxIts;
  return 0;
}

int umt_ctx_get_grey_controls(const umt_ctx* x, double* relTol, int* maxIts, int* enforceHardIterMax) {
  if (!x) return fail(1, "null context");
  if (relTol) *relTol = x->gta.epsPoint;
  if (maxIts) *maxIts = x->gta.maxSweeps;
  if (enforceHardIterMax) *enforceHardIterMax = x->gta.hardIterMax ? 1 : 0;
  return 0;
}

int umt_ctx_set_gta_options(umt_ctx* x, int enforceHardIterMax) {
  if (!x) return fail(1, "null context");
  x->gta.hardIterMax = enforceHardIterMax != 0;
  return 0;
}

int umt_ctx_gta_stats(const umt_ctx* x, int64_t* greyIterations, int64_t* greySweeps) {
  if (!x) return fail(1, "null context");
  if (greyIterations) *greyIterations = x->gta.greyIterations;
  if (greySweeps) *greySweeps = x->gta.greySweeps;
  return 0;
}

}  // extern "C"

namespace {

// Grey opacities for one group (rt/setGTAOpacity.F90:195-233): sigma_s kept
// only above 1e-10 of sigma_t (else sigma_t -> sigma_a), eps = 1/sqrt(1-c)
// so that the stretched problem has no scattering, stretched sigma_t/eps.
// Layout on the device: sigt2 = [grey sigma_t (G,Z) | stretched (G,Z)] in
// sweep precision, eps = [eps (G,Z) | grey sigma_s (G,Z)] in FP64.
int gta_prepare(umt_ctx* x) {
  const size_t GZ = (size_t)x->G * x->Z;
  if (x->sigtHost.size() != GZ || x->sigsHost.size() != GZ)
    return fail(38, "GTA needs sigma_t and sigma_s (umt_ctx_set_sigt, umt_ctx_set_scattering)");
  std::vector<double> st(2 * GZ), ep(2 * GZ);
  for (size_t i = 0; i < GZ; ++i) {
    const double sigt = x->sigtHost[i] + x->tau, sigs = x->sigsHost[i], siga = sigt - sigs;
    double greyScat = sigs, greyTot = sigt;
    if (sigs / sigt <= 1.0e-10) {
      greyScat = 0.0;
      greyTot = siga;
    }
    const double ratio = greyScat / greyTot;
    const double eps = 1.0 / std::sqrt(1.0 - ratio);
    st[i] = greyTot;
    st[GZ + i] = greyTot / eps;
    ep[i] = eps;
    ep[GZ + i] = greyScat;
  }
  GtaState& g = x->gta;
  if (int rc = upload_real(x, g.sigt2, st.data(), 2 * GZ)) return rc;
  HIP_TRY(hipMemcpyAsync(g.eps, ep.data(), 2 * GZ * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  g.luValid = false;  // the transfer matrices follow the opacities
  return 0;
}

// Transfer-matrix solver: T of every zone and group from 8 sweeps of the grey
// directions with no inflow (one per source corner), then the LU factors of
// I - wtiso sigma_s T in place (snac/InitSweepGreyUCBxyz.F90,
// snac/UpdateScalarIntensity.F90:50-97). Tout (may be null) receives T
// before it is factored. The factors stay valid until the opacities change.
template <typename Real>
int gta_build_lu_impl(umt_ctx* x, double* Tout) {
  GtaState& g = x->gta;
  SweepSet& s2 = x->s2;
  const long CG = (long)x->C * x->G, nzg = (long)x->Z * x->G;
  const int A2 = s2.A;
  if (s2.blocks() != 1) return fail(54, "the grey sweeps must hold all their directions in one block");
  if (!g.lu)
    if (int rc = dalloc(&g.lu, 8 * (size_t)CG)) return rc;
  if (!g.localItems) {
    std::vector<int2> it((size_t)x->Z * A2);
    for (int z = 0; z < x->Z; ++z)
      for (int a = 0; a < A2; ++a) it[(size_t)z * A2 + a] = make_int2(z, a);
    if (int rc = dalloc(&g.localItems, it.size())) return rc;
    HIP_TRY(hipMemcpyAsync(g.localItems, it.data(), it.size() * sizeof(int2), hipMemcpyHostToDevice,
                           x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  }
  const void* sigGrey = g.sigt2;
  for (int sc = 0; sc < 8; ++sc) {
    HIP_TRY(umt::launch_gta_unit_source<Real>(sc, x->G, CG, (Real*)g.S2, x->stream));
    umt::SweepLevelArgs<Real> a = level_args<Real>(x, s2, g.S2, sigGrey, 0, 0);
    a.items = g.localItems;
    a.nItems = x->Z * A2;
    a.noInflow = 1;
    HIP_TRY(umt::launch_sweep_level<Real>(a, false, x->stream));
    HIP_TRY(umt::launch_phi_from_psi_acc<Real>((const Real*)s2.psi, s2.weight, A2, x->psiStride(), CG,
                                               g.lu + sc * CG, false, x->stream));
  }
  if (Tout) {
    HIP_TRY(hipMemcpyAsync(Tout, g.lu, 8 * (size_t)CG * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  }
  const size_t GZ = (size_t)x->G * x->Z;
  HIP_TRY(umt::launch_gta_zone_lu(g.lu, g.eps + GZ, x->G, nzg, CG, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  g.luValid = true;
  return 0;
}

int gta_build_lu(umt_ctx* x, double* Tout = nullptr) {
  if (x->gta.luValid && !Tout) return 0;
  return x->precision == 32 ? gta_build_lu_impl<float>(x, Tout) : gta_build_lu_impl<double>(x, Tout);
}

template <typename Real>
int gta_source(umt_ctx* x, int kind, const double* P, const double* other) {
  GtaState& g = x->gta;
  const size_t GZ = (size_t)x->G * x->Z;
  return umt::launch_gta_sweep_source<Real>(kind, P, other, g.eps + GZ, g.eps, x->G,
                                            (long)x->G * x->C, (Real*)g.S2, x->stream) == hipSuccess
             ? 0
             : fail(101, "grey source launch failed");
}

// out = M(P): grey S2 sweep of wtiso (sigma_s P [+ gsrc]) followed by the
// stretched-TSA sweep of wtiso sigma_s eps (out - P) (rt/GreySweep.F90:60-96).
// P == nullptr stands for zero. out must differ from P. With kernel 2,
// residual writes P - M(P) instead (the solve's last step gives it
// directly; see grey_residual).
int grey_sweep(umt_ctx* x, const double* P, bool withSource, double* out, bool residual = false) {
  GtaState& g = x->gta;
  const size_t GZ = (size_t)x->G * x->Z;
  if (g.solver == 2) {
    // One grey sweep of wtiso (sigma_s P [+ gsrc]), then every zone's direct
    // solve: out = P + (I - B)^-1 (phi - P) (rt/GreySweep.F90:12-47).
    if (int rc = gta_build_lu(x)) return rc;
    int rc2 = x->precision == 32 ? gta_source<float>(x, 1, P, withSource ? g.gsrc : nullptr)
                                 : gta_source<double>(x, 1, P, withSource ? g.gsrc : nullptr);
    if (rc2) return rc2;
    if ((rc2 = sweep_set(x, x->s2, g.S2, g.sigt2, out, false, false))) return rc2;
    HIP_TRY(umt::launch_gta_zone_solve(g.lu, P, out, x->G, (long)GZ, (long)x->C * x->G, out,
                                       residual, x->stream));
    g.greySweeps += 1;
    return 0;
  }
  if (residual) return fail(101, "residual form of the grey operator is kernel 2's");
  const char* st = (const char*)g.sigt2;
  const void* sigGrey = st;
  const void* sigStretch = st + GZ * x->rsize();
  int rc;
  if (x->precision == 32) rc = gta_source<float>(x, 1, P, withSource ? g.gsrc : nullptr);
  else rc = gta_source<double>(x, 1, P, withSource ? g.gsrc : nullptr);
  if (rc) return rc;
  if ((rc = sweep_set(x, x->s2, g.S2, sigGrey, out, false, false))) return rc;
  if (x->precision == 32) rc = gta_source<float>(x, 2, P, out);
  else rc = gta_source<double>(x, 2, P, out);
  if (rc) return rc;
  if ((rc = sweep_set(x, x->s2, g.S2, sigStretch, out, true, false))) return rc;
  g.greySweeps += 2;
  return 0;
}

int lincomb(umt_ctx* x, const double* a, const double* b, const double* c, double* out,
            const std::vector<double>& coef) {
  HIP_TRY(hipMemcpyAsync(x->gta.coef, coef.data(), coef.size() * sizeof(double),
                         hipMemcpyHostToDevice, x->stream));
  HIP_TRY(umt::launch_gta_lincomb(a, b, c, x->gta.coef, x->G, (long)x->G * x->C, out, x->stream));
  // coef is host memory reused by the next call: finish this one first.
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int wdot(umt_ctx* x, const double* a, const double* b, std::vector<double>& out) {
  GtaState& g = x->gta;
  const size_t GZ = (size_t)x->G * x->Z;
  HIP_TRY(umt::launch_gta_wdot(a, b, g.eps + GZ, x->vol64, x->G, (long)x->G * x->C, g.dotBlocks,
                               g.partial, g.dots, x->stream));
  out.resize(x->G);
  HIP_TRY(hipMemcpyAsync(out.data(), g.dots, x->G * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

// wn = (x y), wd = (x x) with the sigma_s V weights, in one pass.
int wdot2(umt_ctx* x, const double* a, const double* b, std::vector<double>& wn,
          std::vector<double>& wd) {
  GtaState& g = x->gta;
  const size_t GZ = (size_t)x->G * x->Z;
  HIP_TRY(umt::launch_gta_wdot2(a, b, g.eps + GZ, x->vol64, x->G, (long)x->G * x->C, g.dotBlocks,
                                g.partial, g.dots, x->stream));
  std::vector<double> both(2 * (size_t)x->G);
  HIP_TRY(hipMemcpyAsync(both.data(), g.dots, both.size() * sizeof(double), hipMemcpyDeviceToHost,
                         x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  wn.assign(both.begin(), both.begin() + x->G);
  wd.assign(both.begin() + x->G, both.end());
  return 0;
}

// F += alpha d + omega r; r -= omega As; rr = (r, 1) with the sigma_s V
// weights, in one pass.
int bicg_update(umt_ctx* x, const std::vector<double>& alpha, const std::vector<double>& omega,
                std::vector<double>& rr) {
  GtaState& g = x->gta;
  const size_t GZ = (size_t)x->G * x->Z;
  std::vector<double> cf(3 * (size_t)x->G, 0.0);
  for (int k = 0; k < x->G; ++k) { cf[3 * k] = alpha[k]; cf[3 * k + 1] = omega[k]; }
  HIP_TRY(hipMemcpyAsync(g.coef, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice,
                         x->stream));
  HIP_TRY(umt::launch_gta_bicg_update(g.F, g.d, g.r, g.As, g.coef, g.eps + GZ, x->vol64, x->G,
                                      (long)x->G * x->C, g.dotBlocks, g.partial, g.dots, x->stream));
  rr.resize(x->G);
  HIP_TRY(hipMemcpyAsync(rr.data(), g.dots, x->G * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));  // also before cf (host memory) goes out of scope
  return 0;
}

// out = P - M(P), BiCGStab's operator application (rt/GTASolver.F90:240-245,
// 268-273): kernel 2's zone solve writes it in the same pass, kernel 1 takes
// M(P) off P with a linear combination.
int grey_residual(umt_ctx* x, const double* P, double* out) {
  if (x->gta.solver == 2) return grey_sweep(x, P, false, out, true);
  if (int rc = grey_sweep(x, P, false, out)) return rc;
  std::vector<double> cf(3 * (size_t)x->G);
  for (int k = 0; k < x->G; ++k) { cf[3 * k] = 1.0; cf[3 * k + 1] = -1.0; cf[3 * k + 2] = 0.0; }
  return lincomb(x, P, out, nullptr, out, cf);
}

// BiCGStab on the grey corrections (rt/GTASolver.F90:185-402), the G
// one-group problems in lockstep with their own coefficients. phiHalf is the
// multigroup phi after the sweep; on return g.F holds the correction.
int gta_solve(umt_ctx* x, const double* phiHalf, double epsGrey) {
  GtaState& g = x->gta;
  const int G = x->G;
  const long GC = (long)G * x->C, GZ = (long)G * x->Z;
  const double tiny = 1.0e-300;  // adqtSmall-style floor on inner products
  HIP_TRY(umt::launch_fill_zero<double>(g.F, GC, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>(g.pzOld, GZ, x->stream));
  int rc;
  if ((rc = grey_sweep(x, nullptr, true, g.r))) return rc;
  {
    std::vector<double> cf(3 * G, 0.0);
    for (int k = 0; k < G; ++k) cf[3 * k] = 1.0;
    if ((rc = lincomb(x, g.r, nullptr, nullptr, g.d, cf))) return rc;
  }
  std::vector<double> rrOld, dAd, wn, wd, rr;
  if ((rc = wdot(x, g.r, nullptr, rrOld))) return rc;
  std::vector<char> active(G, 1);
  int nGreyIter = 1;
  for (;;) {
    bool any = false;
    for (int k = 0; k < G; ++k) {
      if (active[k] && std::fabs(rrOld[k]) < tiny) active[k] = 0;
      any = any || active[k];
    }
    if (!any) {
      if (nGreyIter <= 2) {  // no iteration ran: the first residual is the correction
        std::vector<double> cf(3 * G, 0.0);
        for (int k = 0; k < G; ++k) cf[3 * k] = 1.0;
        if ((rc = lincomb(x, g.r, nullptr, nullptr, g.F, cf))) return rc;
      }
      break;
    }
    nGreyIter += 2;
    ++g.greyIterations;
    // Ad = d - M d
    if ((rc = grey_residual(x, g.d, g.Ad))) return rc;
    if ((rc = wdot(x, g.Ad, nullptr, dAd))) return rc;
    std::vector<double> alpha(G, 0.0);
    for (int k = 0; k < G; ++k) {
      if (active[k] && std::fabs(dAd[k]) < tiny) active[k] = 0;
      alpha[k] = active[k] ? rrOld[k] / dAd[k] : 0.0;
    }
    // r = r - alpha Ad
    {
      std::vector<double> cf(3 * G);
      for (int k = 0; k < G; ++k) { cf[3 * k] = 1.0; cf[3 * k + 1] = -alpha[k]; cf[3 * k + 2] = 0.0; }
      if ((rc = lincomb(x, g.r, g.Ad, nullptr, g.r, cf))) return rc;
    }
    // As = r - M r
    if ((rc = grey_residual(x, g.r, g.As))) return rc;
    if ((rc = wdot2(x, g.As, g.r, wn, wd))) return rc;
    std::vector<double> omega(G, 0.0);
    std::vector<char> stopAfter(G, 0);
    for (int k = 0; k < G; ++k) {
      if (!active[k]) continue;
      if (std::fabs(wd[k]) < tiny || std::fabs(wn[k]) < tiny) stopAfter[k] = 1;  // F += alpha d only
      else omega[k] = wn[k] / wd[k];
    }
    // F = F + alpha d + omega r, then r = r - omega As, and r's product
    if ((rc = bicg_update(x, alpha, omega, rr))) return rc;
    // d = r + beta (d - omega Ad)
    {
      std::vector<double> cf(3 * G);
      for (int k = 0; k < G; ++k) {
        const double beta = (active[k] && !stopAfter[k] && omega[k] != 0.0)
                                ? (rr[k] * alpha[k]) / (rrOld[k] * omega[k]) : 0.0;
        cf[3 * k] = beta;
        cf[3 * k + 1] = -beta * omega[k];
        cf[3 * k + 2] = 1.0;
      }
      if ((rc = lincomb(x, g.d, g.Ad, g.r, g.d, cf))) return rc;
    }
    for (int k = 0; k < G; ++k)
      if (stopAfter[k]) active[k] = 0;
    // Convergence of the grey corrections (rt/GTASolver.F90:318-370).
    HIP_TRY(umt::launch_gta_zone_error(g.F, phiHalf, x->vol64, x->volZ, G, GZ, g.pzOld, g.errBlocks,
                                       g.partial, g.err, x->stream));
    double e[3];
    HIP_TRY(hipMemcpyAsync(e, g.err, sizeof e, hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    if (!std::isfinite(e[0]) || !std::isfinite(e[1]) || !std::isfinite(e[2]))
      return fail(39, "grey solver produced a NaN at grey iteration %d", nGreyIter);
    const double relL2 = e[2] != 0.0 ? std::sqrt(std::fabs(e[1] / e[2])) : 0.0;
    const double errGrey = std::max(e[0], relL2);
    if (g.hardIterMax && nGreyIter >= g.maxSweeps) break;  // rt/GTASolver.F90:373-375
    if ((errGrey < g.epsPoint || nGreyIter >= g.maxSweeps) && errGrey < epsGrey) break;
    if (nGreyIter >= 100 * g.maxSweeps)
      return fail(40, "grey solver not converging after %d grey iterations (error %.3e)", nGreyIter, errGrey);
    rrOld = rr;
  }
  return 0;
}

// Host-array wrappers of the grey operator and solve (their parity is
// checked piece by piece against the oracle).
int grey_upload(umt_ctx* x, double* dst, const double* src) {
  const size_t GC = (size_t)x->G * x->C;
  if (src) HIP_TRY(hipMemcpyAsync(dst, src, GC * sizeof(double), hipMemcpyHostToDevice, x->stream));
  else HIP_TRY(umt::launch_fill_zero<double>(dst, (long)GC, x->stream));
  return 0;
}

int learn_prepare(umt_ctx* x) {
  LearnState& L = x->learn;
  if (!L.zoneNbr) return fail(39, "learned acceleration needs umt_ctx_set_zone_neighbors first");
  if (!x->haveSigt) return fail(40, "learned acceleration needs umt_ctx_set_sigt first");
  const size_t GZ = (size_t)x->G * x->Z;
  int rc = 0;
  if (!L.pp) {
    if ((rc = dalloc(&L.sigt64, GZ)) || (rc = dalloc(&L.pp, GZ)) || (rc = dalloc(&L.dp, GZ)) ||
        (rc = dalloc(&L.rel, GZ)) || (rc = dalloc(&L.ratio, GZ)) || (rc = dalloc(&L.ta, GZ)) ||
        (rc = dalloc(&L.tb, GZ)) || (rc = dalloc(&L.tpart, (size_t)umt::kLearnTargetBlocks)) ||
        (rc = dalloc(&L.maxRel, (size_t)x->G)) || (rc = dalloc(&L.scratchBits, 2)) ||
        (rc = dalloc(&L.X, GZ * umt::kMlpFeatures)) || (rc = dalloc(&L.r, GZ)))
      return rc;
  }
  std::vector<double> st(x->sigtHost);
  for (double& v : st) v += x->tau;
  HIP_TRY(hipMemcpyAsync(L.sigt64, st.data(), GZ * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(umt::launch_fill_zero<double>((double*)L.scratchBits, 2, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

// Keeps one plain state (features, zone averages, change) for training.
int learn_record(umt_ctx* x, double res) {
  LearnState& L = x->learn;
  const size_t GZ = (size_t)x->G * x->Z;
  LearnState::Pending p{nullptr, nullptr, nullptr, res};
  int rc;
  if ((rc = dalloc(&p.X, GZ * umt::kMlpFeatures)) || (rc = dalloc(&p.ph, GZ)) || (rc = dalloc(&p.d, GZ))) {
    dfree(p.X); dfree(p.ph); dfree(p.d);
    return rc;
  }
  L.pending.push_back(p);
  HIP_TRY(hipMemcpyAsync(p.X, L.X, GZ * umt::kMlpFeatures * sizeof(float), hipMemcpyDeviceToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(p.ph, x->phibar, GZ * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
  HIP_TRY(umt::launch_learn_delta(x->phibar, L.pp, (long)GZ, p.d, x->stream));
  return 0;
}

}  // namespace

extern "C" {

int umt_ctx_autotune_grey_layouts(umt_ctx* x, int reps, int* bestK, double* msPerLayout) {
  if (!x || reps < 1) return fail(1, "bad argument");
  if (!x->gta.enabled) return fail(37, "GTA needs umt_ctx_enable_gta first");
  if (int rc = gta_prepare(x)) return rc;
  GtaState& g = x->gta;
  const long GC = (long)x->G * x->C;
  if (x->precision == 32) HIP_TRY(umt::launch_fill_zero<float>((float*)g.S2, GC, x->stream));
  else HIP_TRY(umt::launch_fill_zero<double>((double*)g.S2, GC, x->stream));
  static const int cand[5] = {1, 2, 4, 8, umt::kZoneLayout};
  double best = 1e300;
  int bk = x->s2.groupsPerLane;
  for (int i = 0; i < 5; ++i) {
    if (msPerLayout) msPerLayout[i] = -1.0;
    if (!umt::sweep_layout_supported(x->G, cand[i])) continue;
    x->s2.groupsPerLane = cand[i];
    if (int rc = sweep_set(x, x->s2, g.S2, g.sigt2, g.As, false, false)) return rc;  // warm-up
    HIP_TRY(hipEventRecord(x->ev0, x->stream));
    for (int r = 0; r < reps; ++r)
      if (int rc = sweep_set(x, x->s2, g.S2, g.sigt2, g.As, false, false)) return rc;
    HIP_TRY(hipEventRecord(x->ev1, x->stream));
    HIP_TRY(hipEventSynchronize(x->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, x->ev0, x->ev1));
    if (msPerLayout) msPerLayout[i] = t / reps;
    if (t / reps < best) {
      best = t / reps;
      bk = cand[i];
    }
  }
  x->s2.groupsPerLane = bk;
  if (bestK) *bestK = bk;
  return 0;
}

int umt_ctx_autotune_grey(umt_ctx* x, int reps, int* bestK, double* msPerK) {
  double ms[5];
  if (int rc = umt_ctx_autotune_grey_layouts(x, reps, bestK, ms)) return rc;
  if (msPerK)
    for (int i = 0; i < 4; ++i) msPerK[i] = ms[i];
  return 0;
}

int umt_ctx_grey_sweep(umt_ctx* x, const double* P, const double* gsrc, double* out) {
  if (!x || !out) return fail(1, "null argument");
  if (!x->gta.enabled) return fail(37, "GTA needs umt_ctx_enable_gta first");
  if (int rc = gta_prepare(x)) return rc;
  GtaState& g = x->gta;
  const size_t GC = (size_t)x->G * x->C;
  int rc;
  if ((rc = grey_upload(x, g.P1, P)) || (rc = grey_upload(x, g.gsrc, gsrc))) return rc;
  if ((rc = grey_sweep(x, P ? g.P1 : nullptr, gsrc != nullptr, g.As))) return rc;
  HIP_TRY(hipMemcpyAsync(out, g.As, GC * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_grey_transfer(umt_ctx* x, double* T, double* LU) {
  if (!x || !T) return fail(1, "null argument");
  if (!x->gta.enabled) return fail(37, "GTA needs umt_ctx_enable_gta first");
  if (int rc = gta_prepare(x)) return rc;
  if (int rc = gta_build_lu(x, T)) return rc;
  if (LU) {
    HIP_TRY(hipMemcpyAsync(LU, x->gta.lu, 8 * (size_t)x->C * x->G * sizeof(double), hipMemcpyDeviceToHost,
                           x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
  }
  return 0;
}

int umt_ctx_grey_solve(umt_ctx* x, const double* gsrc, const double* phiHalf, double epsGrey,
                       double* F, int64_t* greyIterations) {
  if (!x || !gsrc || !phiHalf || !F) return fail(1, "null argument");
  if (!x->gta.enabled) return fail(37, "GTA needs umt_ctx_enable_gta first");
  if (int rc = gta_prepare(x)) return rc;
  GtaState& g = x->gta;
  const size_t GC = (size_t)x->G * x->C;
  int rc;
  if ((rc = grey_upload(x, g.gsrc, gsrc)) || (rc = grey_upload(x, g.phiHalf, phiHalf))) return rc;
  const long long before = g.greyIterations;
  if ((rc = gta_solve(x, g.phiHalf, epsGrey))) return rc;
  HIP_TRY(hipMemcpyAsync(F, g.F, GC * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  if (greyIterations) *greyIterations = g.greyIterations - before;
  return 0;
}

int umt_ctx_source_iteration(umt_ctx* x, double tol, int maxIters, int* iters, double* residual,
                             double* hist) {
  if (!x) return fail(1, "null context");
  if (!x->haveScattering) return fail(33, "source iteration needs umt_ctx_set_scattering first");
  // sigma_a = sigma_t - sigma_s >= 0: with more scattering than collision
  // the iteration diverges (and the grey opacities have no real stretch).
  for (size_t i = 0; i < x->sigsHost.size() && i < x->sigtHost.size(); ++i)
    if (x->sigsHost[i] > x->sigtHost[i])
      return fail(52, "sigs entry %zu exceeds Sigt (scattering ratio above 1)", i + 1);
  const long nGC = (long)x->G * x->C, nGZ = (long)x->G * x->Z;
  double res = 1.0;
  int it = 0;
  // Zone averages of the starting phi are the "old" values of iteration 1.
  HIP_TRY(umt::launch_residual(x->phi, x->vol64, x->volZ, x->G, nGZ, x->phibar, x->phibarOld, 1e-6,
                               x->bits, x->stream));
  const bool gta = x->acceleration == 1 || x->acceleration == 3;
  const bool learned = x->acceleration == 2 || x->acceleration == 3;
  // Learned-correction bookkeeping (corrections, or recording states).
  const bool track = learned || x->learn.collect;
  if (gta)
    if (int rc = gta_prepare(x)) return rc;
  if (track)
    if (int rc = learn_prepare(x)) return rc;
  LearnState& L = x->learn;
  bool correcting = learned, haveDp = false;
  int since = 0, eligible = 0;
  double lastCorrectedRes = -1.0;
  L.fellBack = 0;
  double epsGrey = x->gta.epsGreyInit;
  double bestRes = HUGE_VAL;
  int bestIt = 0;
  bool stalled = false;
  while (it < maxIters) {
    if (gta || track)
      HIP_TRY(hipMemcpyAsync(x->phiPrev, x->phi, nGC * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
    if (track)
      HIP_TRY(hipMemcpyAsync(L.pp, x->phibarOld, nGZ * sizeof(double), hipMemcpyDeviceToDevice, x->stream));
    if (x->precision == 32)
      HIP_TRY(umt::launch_source_update<float>(x->phi, x->sigs, x->q, x->G, nGC, (float*)x->S, x->stream));
    else
      HIP_TRY(umt::launch_source_update<double>(x->phi, x->sigs, x->q, x->G, nGC, (double*)x->S, x->stream));
    if (int rc = sweep(x, false)) return rc;
    if (gta) {
      // Grey source from the scattering-rate change, solve, add the
      // correction (rt/LinearSolver.F90:81-111, rt/addGreyCorrections.F90).
      HIP_TRY(umt::launch_gta_grey_source(x->phi, x->phiPrev, x->sigs, x->G, nGC, x->gta.gsrc, x->stream));
      if (int rc = gta_solve(x, x->phi, epsGrey)) return rc;
      std::vector<double> cf(3 * (size_t)x->G);
      for (int k = 0; k < x->G; ++k) { cf[3 * k] = 1.0; cf[3 * k + 1] = 1.0; cf[3 * k + 2] = 0.0; }
      if (int rc = lincomb(x, x->phi, x->gta.F, nullptr, x->phi, cf)) return rc;
    }
    // Corrected iterates may go negative; their residual tests |phibar|.
    HIP_TRY(umt::launch_residual(x->phi, x->vol64, x->volZ, x->G, nGZ, x->phibar, x->phibarOld, 1e-6,
                                 x->bits, x->stream, learned));
    if (track)  // d = phibar - pp, the network's features (X) and the previous change dp
      HIP_TRY(umt::launch_learn_features(x->phibar, L.pp, haveDp ? L.dp : nullptr, L.sigt64, x->sigs,
                                         x->q, L.zoneNbr, x->G, nGZ, L.rel, L.ratio, L.maxRel, L.X,
                                         x->stream));
    unsigned long long b[2], badFlag = 0;
    HIP_TRY(hipMemcpyAsync(b, x->bits, sizeof b, hipMemcpyDeviceToHost, x->stream));
    if (track && L.scratchBits)
      HIP_TRY(hipMemcpyAsync(&badFlag, L.scratchBits + 1, sizeof badFlag, hipMemcpyDeviceToHost, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    if (badFlag && correcting) {
      // A correction came out non-finite (those entries were left uncorrected):
      // stop correcting, plain iteration carries on.
      correcting = false;
      L.fellBack = it + 1;
    }
    double r;
    std::memcpy(&r, &b[1], sizeof r);
    if (!std::isfinite(r)) return fail(34, "source iteration residual is not finite at iteration %d", it + 1);
    res = r;
    if (hist) hist[it] = r;
    ++it;
    epsGrey = res / 20.0;  // rt/rtmainsn.F90:220
    if (res < tol) break;
    if (res < bestRes) {
      bestRes = res;
      bestIt = it;
    } else if (x->stallWindow > 0 && it - bestIt >= x->stallWindow) {
      stalled = true;  // e.g. at the rounding floor of FP32 sweeps
      break;
    }
    if (track) {
      // Plain source iteration arbitrates every correction: once a corrected
      // state's sweep changes phi more than the sweep the correction
      // followed, corrections stop and the iteration continues plain.
      if (lastCorrectedRes >= 0.0) {
        if (L.guard && res > lastCorrectedRes && correcting) {
          correcting = false;
          L.fellBack = it;
        }
        lastCorrectedRes = -1.0;
      }
      ++since;
      if (since >= L.every && haveDp) {
        if (L.collect && eligible++ % L.collectStride == 0)
          if (int rc = learn_record(x, res)) return rc;
        if (correcting) {
          HIP_TRY(umt::launch_mlp_forward(L.X, nGZ, L.net->packed, L.net->params, L.r, x->stream));
          HIP_TRY(umt::launch_learn_apply(x->phi, x->phiPrev, L.r, x->G, nGC, L.scratchBits + 1,
                                          x->stream));
          // The corrected state's zone averages are the "old" values of the next residual.
          HIP_TRY(umt::launch_zone_average(x->phi, x->vol64, x->volZ, x->G, nGZ, x->phibarOld,
                                           L.scratchBits, x->stream));
          since = 0;
          lastCorrectedRes = res;
          ++L.corrections;
        }
      }
      HIP_TRY(umt::launch_learn_delta(x->phibar, L.pp, nGZ, L.dp, x->stream));
      haveDp = true;
    }
  }
  if (iters) *iters = it;
  if (residual) *residual = res;
  return res < tol ? 0 : (stalled ? 2 : 1);
}

int umt_ctx_save_psi(umt_ctx* x) {
  if (!x) return fail(1, "null context");
  if (!x->psiPrev) return fail(49, "umt_ctx_set_tau with tau > 0 first");
  if (!x->haveScattering) return fail(33, "umt_ctx_save_psi needs umt_ctx_set_scattering first");
  // The final sweep of a cycle (rt/rtmainsn.F90:238-246): source from the
  // converged phi, sweep, Psi = psi of every angle. The tau Psi source of
  // this sweep still reads the old Psi, so psi goes to a scratch copy first.
  const long nGC = (long)x->G * x->C;
  const size_t n = (size_t)x->psiStride() * x->sn.A;
  void* fresh = nullptr;
  if (int rc = dalloc((char**)&fresh, n * x->rsize())) return rc;
  int rc = 0;
  if (x->precision == 32) {
    if (umt::launch_source_update<float>(x->phi, x->sigs, x->q, x->G, nGC, (float*)x->S, x->stream) != hipSuccess)
      rc = fail(101, "source update failed");
    if (!rc) rc = launch_sweep<float>(x, x->sn, x->S, x->sigt, x->phi, false, false, fresh);
  } else {
    if (umt::launch_source_update<double>(x->phi, x->sigs, x->q, x->G, nGC, (double*)x->S, x->stream) != hipSuccess)
      rc = fail(101, "source update failed");
    if (!rc) rc = launch_sweep<double>(x, x->sn, x->S, x->sigt, x->phi, false, false, fresh);
  }
  if (!rc && hipMemcpyAsync(x->psiPrev, fresh, n * x->rsize(), hipMemcpyDeviceToDevice, x->stream) != hipSuccess)
    rc = fail(101, "psi copy failed");
  if (!rc && hipStreamSynchronize(x->stream) != hipSuccess) rc = fail(101, "synchronize failed");
  (void)hipStreamSynchronize(x->stream);
  dfree(fresh);
  return rc;
}

int umt_ctx_set_stall_window(umt_ctx* x, int window) {
  if (!x) return fail(1, "null context");
  if (window < 0) return fail(1, "stall window must be >= 0");
  x->stallWindow = window;
  return 0;
}

int umt_ctx_time_sweeps(umt_ctx* x, int warmup, int reps, double* msSweep, double* msLevels,
                        int* launches) {
  if (!x || reps < 1) return fail(1, "bad argument");
  for (int i = 0; i < warmup; ++i)
    if (int rc = sweep(x, false)) return rc;
  HIP_TRY(hipStreamSynchronize(x->stream));
  const int L = (int)x->sn.levelOffset.size() - 1;
  double sumLevels = 0.0;
  HIP_TRY(hipEventRecord(x->ev0, x->stream));
  for (int r = 0; r < reps; ++r) {
    if (int rc = sweep(x, true)) return rc;
    HIP_TRY(hipStreamSynchronize(x->stream));
    for (int h = 0; h < L; ++h) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, x->sn.evLevel[2 * h], x->sn.evLevel[2 * h + 1]));
      sumLevels += ms;
    }
  }
  HIP_TRY(hipEventRecord(x->ev1, x->stream));
  HIP_TRY(hipEventSynchronize(x->ev1));
  float total = 0.f;
  HIP_TRY(hipEventElapsedTime(&total, x->ev0, x->ev1));
  if (msSweep) *msSweep = total / reps;
  if (msLevels) *msLevels = sumLevels / reps;
  if (launches) *launches = L;
  return 0;
}

// ---------------------------------------------------------------------------
// Learned correction network

}  // extern "C"

namespace {

constexpr float kAdamBeta1 = 0.9f, kAdamBeta2 = 0.999f, kAdamEps = 1e-8f;
constexpr int kTrainStepsPerWave = 1;  // 32 samples per wave: a batch spreads over more CUs

// splitmix64 stream of uniform doubles in [0, 1).
double next_unit(uint64_t& state) {
  state += 0x9E3779B97F4A7C15ull;
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

int learner_upload_params(umt_learner* L, const float* host) {
  HIP_TRY(hipMemcpyAsync(L->params, host, umt::kMlpParams * sizeof(float), hipMemcpyHostToDevice, L->stream));
  HIP_TRY(umt::launch_fill_zero<float>(L->m, umt::kMlpParams, L->stream));
  HIP_TRY(umt::launch_fill_zero<float>(L->v, umt::kMlpParams, L->stream));
  HIP_TRY(umt::launch_mlp_pack(L->params, L->packed, L->stream));
  HIP_TRY(hipStreamSynchronize(L->stream));
  L->adamSteps = 0;
  return 0;
}

// Room for `need` training samples, keeping the ones held.
int learner_reserve(umt_learner* L, long need) {
  if (need <= L->cap) return 0;
  long cap = std::max(need, 2 * L->cap);
  float *X = nullptr, *a = nullptr, *b = nullptr;
  int rc;
  if ((rc = dalloc(&X, (size_t)cap * umt::kMlpFeatures)) || (rc = dalloc(&a, (size_t)cap)) ||
      (rc = dalloc(&b, (size_t)cap))) {
    dfree(X); dfree(a); dfree(b);
    return rc;
  }
  HIP_TRY(hipStreamSynchronize(L->stream));
  if (L->n) {
    HIP_TRY(hipMemcpy(X, L->X, (size_t)L->n * umt::kMlpFeatures * sizeof(float), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(a, L->a, (size_t)L->n * sizeof(float), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(b, L->b, (size_t)L->n * sizeof(float), hipMemcpyDeviceToDevice));
  }
  dfree(L->X); dfree(L->a); dfree(L->b);
  L->X = X; L->a = a; L->b = b; L->cap = cap;
  return 0;
}

int learner_partial(umt_learner* L, long rows) {
  if (rows <= L->partialCap) return 0;
  HIP_TRY(hipStreamSynchronize(L->stream));  // earlier steps may still read it
  dfree(L->partial);
  if (int rc = dalloc(&L->partial, (size_t)rows * (umt::kMlpParams + 1))) return rc;
  L->partialCap = rows;
  return 0;
}

// Gradient (and, with lr > 0, one Adam step) over samples [start, start +
// count) of (X, a, b), in idx order when idx is given.
int learner_step(umt_learner* L, const float* X, const float* a, const float* b, const int* idx,
                 long start, long count, float lr) {
  const int waves = umt::mlp_train_waves(count, kTrainStepsPerWave);
  if (int rc = learner_partial(L, waves)) return rc;
  HIP_TRY(umt::launch_mlp_train(X, a, b, idx, start, count, kTrainStepsPerWave, L->packed, L->params,
                                L->partial, L->stream));
  float bc1 = 1.f, bc2 = 1.f;
  if (lr > 0.f) {
    ++L->adamSteps;
    bc1 = (float)(1.0 - std::pow((double)kAdamBeta1, (double)L->adamSteps));
    bc2 = (float)(1.0 - std::pow((double)kAdamBeta2, (double)L->adamSteps));
  }
  HIP_TRY(umt::launch_mlp_reduce_adam(L->partial, waves, L->grad, L->params, L->m, L->v, lr,
                                      kAdamBeta1, kAdamBeta2, kAdamEps, bc1, bc2, L->stream));
  if (lr > 0.f) HIP_TRY(umt::launch_mlp_pack(L->params, L->packed, L->stream));
  return 0;
}

// Host arrays of n samples uploaded to temporary device buffers.
struct SampleUpload {
  float *X = nullptr, *a = nullptr, *b = nullptr;
  ~SampleUpload() { dfree(X); dfree(a); dfree(b); }
  int load(hipStream_t s, const float* hX, const float* ha, const float* hb, long n) {
    int rc;
    if ((rc = dalloc(&X, (size_t)n * umt::kMlpFeatures))) return rc;
    HIP_TRY(hipMemcpyAsync(X, hX, (size_t)n * umt::kMlpFeatures * sizeof(float), hipMemcpyHostToDevice, s));
    if (ha) {
      if ((rc = dalloc(&a, (size_t)n)) || (rc = dalloc(&b, (size_t)n))) return rc;
      HIP_TRY(hipMemcpyAsync(a, ha, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(b, hb, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    }
    return 0;
  }
};

}  // namespace

extern "C" {

int umt_learner_num_params(void) { return umt::kMlpParams; }

void umt_learner_destroy(umt_learner* L) {
  if (!L) return;
  if (L->stream) (void)hipStreamSynchronize(L->stream);
  dfree(L->params); dfree(L->m); dfree(L->v); dfree(L->grad); dfree(L->packed);
  dfree(L->X); dfree(L->a); dfree(L->b); dfree(L->perm); dfree(L->partial); dfree(L->lossLog);
  if (L->stream) (void)hipStreamDestroy(L->stream);
  delete L;
}

int umt_learner_create(uint64_t seed, umt_learner** out) {
  if (!out) return fail(1, "null output pointer");
  *out = nullptr;
  umt_learner* L = new umt_learner();
  auto bail = [&](int rc) {
    std::string keep = g_err;
    umt_learner_destroy(L);
    g_err = keep;
    return rc;
  };
  int rc;
  if (hipStreamCreate(&L->stream) != hipSuccess) return bail(fail(2, "hipStreamCreate failed"));
  if ((rc = dalloc(&L->params, umt::kMlpParams)) || (rc = dalloc(&L->m, umt::kMlpParams)) ||
      (rc = dalloc(&L->v, umt::kMlpParams)) || (rc = dalloc(&L->grad, umt::kMlpParams + 1)) ||
      (rc = dalloc(&L->packed, umt::kMlpPacked)))
    return bail(rc);
  // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases; the
  // feature slots beyond the used ones get zero weights.
  std::vector<float> p(umt::kMlpParams, 0.f);
  uint64_t st = seed;
  auto draw = [&](double bound) { return (float)((2.0 * next_unit(st) - 1.0) * bound); };
  const double b1 = 1.0 / std::sqrt((double)umt::kMlpUsedFeatures), b2 = 1.0 / std::sqrt((double)umt::kMlpHidden);
  for (int f = 0; f < umt::kMlpFeatures; ++f)
    for (int h = 0; h < umt::kMlpHidden; ++h)
      p[umt::kMlpW1 + f * umt::kMlpHidden + h] = f < umt::kMlpUsedFeatures ? draw(b1) : 0.f;
  for (int h = 0; h < umt::kMlpHidden; ++h) p[umt::kMlpB1 + h] = draw(b1);
  for (int i = 0; i < umt::kMlpHidden * umt::kMlpHidden; ++i) p[umt::kMlpW2 + i] = draw(b2);
  for (int h = 0; h < umt::kMlpHidden; ++h) p[umt::kMlpB2 + h] = draw(b2);
  for (int h = 0; h < umt::kMlpHidden; ++h) p[umt::kMlpW3 + h] = draw(b2);
  p[umt::kMlpB3] = draw(b2);
  if ((rc = learner_upload_params(L, p.data()))) return bail(rc);
  *out = L;
  return 0;
}

int umt_learner_get_params(const umt_learner* L, float* params) {
  if (!L || !params) return fail(1, "null argument");
  HIP_TRY(hipMemcpyAsync(params, L->params, umt::kMlpParams * sizeof(float), hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(hipStreamSynchronize(L->stream));
  return 0;
}

int umt_learner_set_params(umt_learner* L, const float* params) {
  if (!L || !params) return fail(1, "null argument");
  for (int i = 0; i < umt::kMlpParams; ++i)
    if (!std::isfinite(params[i])) return fail(41, "parameter %d is not finite", i);
  return learner_upload_params(L, params);
}

int umt_learner_forward(umt_learner* L, const float* X, int64_t n, float* r) {
  if (!L || (n > 0 && (!X || !r)) || n < 0) return fail(1, "bad argument");
  if (n == 0) return 0;
  SampleUpload u;
  if (int rc = u.load(L->stream, X, nullptr, nullptr, n)) return rc;
  float* dr = nullptr;
  if (int rc = dalloc(&dr, (size_t)n)) return rc;
  hipError_t e = umt::launch_mlp_forward(u.X, n, L->packed, L->params, dr, L->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(r, dr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, L->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  dfree(dr);
  HIP_TRY(e);
  return 0;
}

int umt_learner_gradient(umt_learner* L, const float* X, const float* a, const float* b, int64_t n,
                         float* grad, double* sse) {
  if (!L || !X || !a || !b || n < 1) return fail(1, "bad argument");
  SampleUpload u;
  if (int rc = u.load(L->stream, X, a, b, n)) return rc;
  if (int rc = learner_step(L, u.X, u.a, u.b, nullptr, 0, n, 0.f)) return rc;
  std::vector<float> g(umt::kMlpParams + 1);
  HIP_TRY(hipMemcpyAsync(g.data(), L->grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(hipStreamSynchronize(L->stream));
  if (grad) std::memcpy(grad, g.data(), umt::kMlpParams * sizeof(float));
  if (sse) *sse = g[umt::kMlpParams];
  return 0;
}

int umt_learner_step(umt_learner* L, const float* X, const float* a, const float* b, int64_t n,
                     double lr) {
  if (!L || !X || !a || !b || n < 1 || !(lr > 0.0)) return fail(1, "bad argument");
  SampleUpload u;
  if (int rc = u.load(L->stream, X, a, b, n)) return rc;
  if (int rc = learner_step(L, u.X, u.a, u.b, nullptr, 0, n, (float)lr)) return rc;
  HIP_TRY(hipStreamSynchronize(L->stream));
  return 0;
}

int umt_learner_add_samples(umt_learner* L, const float* X, const float* a, const float* b, int64_t n) {
  if (!L || n < 0 || (n > 0 && (!X || !a || !b))) return fail(1, "bad argument");
  if (n == 0) return 0;
  if (int rc = learner_reserve(L, L->n + n)) return rc;
  HIP_TRY(hipMemcpyAsync(L->X + (size_t)L->n * umt::kMlpFeatures, X, (size_t)n * umt::kMlpFeatures * sizeof(float),
                         hipMemcpyHostToDevice, L->stream));
  HIP_TRY(hipMemcpyAsync(L->a + L->n, a, (size_t)n * sizeof(float), hipMemcpyHostToDevice, L->stream));
  HIP_TRY(hipMemcpyAsync(L->b + L->n, b, (size_t)n * sizeof(float), hipMemcpyHostToDevice, L->stream));
  HIP_TRY(hipStreamSynchronize(L->stream));
  L->n += n;
  return 0;
}

int umt_learner_num_samples(const umt_learner* L, int64_t* n) {
  if (!L || !n) return fail(1, "null argument");
  *n = L->n;
  return 0;
}

int umt_learner_get_samples(const umt_learner* L, float* X, float* a, float* b) {
  if (!L || !X || !a || !b) return fail(1, "null argument");
  if (!L->n) return 0;
  HIP_TRY(hipMemcpyAsync(X, L->X, (size_t)L->n * umt::kMlpFeatures * sizeof(float), hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(hipMemcpyAsync(a, L->a, (size_t)L->n * sizeof(float), hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(hipMemcpyAsync(b, L->b, (size_t)L->n * sizeof(float), hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(hipStreamSynchronize(L->stream));
  return 0;
}

int umt_learner_clear_samples(umt_learner* L) {
  if (!L) return fail(1, "null learner");
  L->n = 0;
  return 0;
}

int umt_learner_train(umt_learner* L, int epochs, int batch, double lr, uint64_t seed, double* meanLoss) {
  if (!L || epochs < 1 || batch < 1 || !(lr > 0.0)) return fail(1, "bad argument");
  if (L->n < 1) return fail(42, "the learner holds no training samples");
  const long n = L->n;
  const long B = std::min<long>(batch, n);
  const long steps = (n + B - 1) / B;
  if (n > L->permCap) {
    dfree(L->perm);
    if (int rc = dalloc(&L->perm, (size_t)n)) return rc;
    L->permCap = n;
  }
  if (steps > L->lossCap) {
    dfree(L->lossLog);
    if (int rc = dalloc(&L->lossLog, (size_t)steps)) return rc;
    L->lossCap = steps;
  }
  std::vector<int> perm(n);
  uint64_t st = seed;
  for (int e = 0; e < epochs; ++e) {
    for (long i = 0; i < n; ++i) perm[i] = (int)i;
    for (long i = n - 1; i > 0; --i) {  // Fisher-Yates
      const long j = (long)(next_unit(st) * (double)(i + 1));
      std::swap(perm[i], perm[std::min(j, i)]);
    }
    HIP_TRY(hipMemcpyAsync(L->perm, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, L->stream));
    for (long k = 0; k < steps; ++k) {
      const long start = k * B, count = std::min(B, n - start);
      if (int rc = learner_step(L, L->X, L->a, L->b, L->perm, start, count, (float)lr)) return rc;
      HIP_TRY(hipMemcpyAsync(L->lossLog + k, L->grad + umt::kMlpParams, sizeof(float),
                             hipMemcpyDeviceToDevice, L->stream));
    }
    HIP_TRY(hipStreamSynchronize(L->stream));  // perm is rewritten next epoch
  }
  std::vector<float> loss(steps);
  HIP_TRY(hipMemcpyAsync(loss.data(), L->lossLog, steps * sizeof(float), hipMemcpyDeviceToHost, L->stream));
  HIP_TRY(hipStreamSynchronize(L->stream));
  double s = 0.0;
  for (float v : loss) s += v;
  if (meanLoss) *meanLoss = s / (double)n;
  if (!std::isfinite(s)) return fail(43, "training loss is not finite");
  return 0;
}

int umt_ctx_set_zone_neighbors(umt_ctx* x, const int* zoneOpp) {
  if (!x || !zoneOpp) return fail(1, "null argument");
  std::vector<int> nb(6 * (size_t)x->Z);
  for (size_t i = 0; i < nb.size(); ++i) {
    const int z = zoneOpp[i];
    if (z > x->Z) return fail(44, "zoneOpp entry %zu (%d) exceeds the zone count", i + 1, z);
    nb[i] = z > 0 ? z - 1 : -1;
  }
  if (!x->learn.zoneNbr)
    if (int rc = dalloc(&x->learn.zoneNbr, nb.size())) return rc;
  HIP_TRY(hipMemcpyAsync(x->learn.zoneNbr, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_set_learner(umt_ctx* x, umt_learner* L, int every, int guard) {
  if (!x) return fail(1, "null context");
  if (every < 1) return fail(1, "every must be at least 1");
  x->learn.net = L;
  x->learn.every = every;
  x->learn.guard = guard != 0;
  if (!L && x->acceleration >= 2) x->acceleration = x->acceleration == 3 ? 1 : 0;
  return 0;
}

int umt_ctx_learn_collect(umt_ctx* x, int on, int stride) {
  if (!x) return fail(1, "null context");
  if (stride < 1) return fail(1, "collection stride must be at least 1");
  x->learn.collect = on != 0;
  x->learn.collectStride = stride;
  return 0;
}

int umt_ctx_learn_commit(umt_ctx* x, umt_learner* L, double minResidual, int* nStates) {
  if (!x || !L) return fail(1, "null argument");
  LearnState& S = x->learn;
  const long GZ = (long)x->G * x->Z;
  HIP_TRY(hipStreamSynchronize(x->stream));
  int kept = 0;
  for (auto& p : S.pending)
    if (p.res >= minResidual) ++kept;
  if (int rc = learner_reserve(L, L->n + (long)kept * GZ)) return rc;
  HIP_TRY(hipStreamSynchronize(L->stream));
  // Targets against the current zone averages (the converged state).
  for (auto& p : S.pending) {
    if (p.res < minResidual) continue;
    HIP_TRY(umt::launch_learn_targets(p.ph, p.d, x->phibar, GZ, S.ta, S.tb, S.tpart, L->a + L->n,
                                      L->b + L->n, x->stream));
    HIP_TRY(hipMemcpyAsync(L->X + (size_t)L->n * umt::kMlpFeatures, p.X, (size_t)GZ * umt::kMlpFeatures * sizeof(float),
                           hipMemcpyDeviceToDevice, x->stream));
    L->n += GZ;
  }
  HIP_TRY(hipStreamSynchronize(x->stream));
  free_pending(S);
  if (nStates) *nStates = kept;
  return 0;
}

int umt_ctx_learn_discard(umt_ctx* x) {
  if (!x) return fail(1, "null context");
  HIP_TRY(hipStreamSynchronize(x->stream));
  free_pending(x->learn);
  return 0;
}

int umt_ctx_learn_features(umt_ctx* x, const double* ph, const double* pp, const double* dp, float* X) {
  if (!x || !ph || !pp || !X) return fail(1, "null argument");
  if (int rc = learn_prepare(x)) return rc;
  LearnState& L = x->learn;
  const size_t GZ = (size_t)x->G * x->Z;
  HIP_TRY(hipMemcpyAsync(L.ta, ph, GZ * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(hipMemcpyAsync(L.pp, pp, GZ * sizeof(double), hipMemcpyHostToDevice, x->stream));
  if (dp) HIP_TRY(hipMemcpyAsync(L.dp, dp, GZ * sizeof(double), hipMemcpyHostToDevice, x->stream));
  HIP_TRY(umt::launch_learn_features(L.ta, L.pp, dp ? L.dp : nullptr, L.sigt64, x->sigs, x->q,
                                     L.zoneNbr, x->G, (long)GZ, L.rel, L.ratio, L.maxRel, L.X, x->stream));
  HIP_TRY(hipMemcpyAsync(X, L.X, GZ * umt::kMlpFeatures * sizeof(float), hipMemcpyDeviceToHost, x->stream));
  HIP_TRY(hipStreamSynchronize(x->stream));
  return 0;
}

int umt_ctx_learn_stats(const umt_ctx* x, int64_t* corrections, int* fellBackAt) {
  if (!x) return fail(1, "null context");
  if (corrections) *corrections = x->learn.corrections;
  if (fellBackAt) *fellBackAt = x->learn.fellBack;
  return 0;
}

// ---------------------------------------------------------------------------
// Drop-in per-angle sweep

namespace {

// Device buffers of one drop-in call: the host arrays' device copies.
// Corner arrays are (3, maxcf, C) and (maxcf, C) in size; the hex kernels
// take them zone-blocked with maxcf = 3, the polyhedral kernel in Teton's
// layout, with cEZ, nCFaces, nextC, numCorner and cOffSet beside them.
struct DropBufs {
  double *Afp = nullptr, *Aez = nullptr, *vol = nullptr, *S = nullptr, *sigt = nullptr,
         *omega = nullptr, *psi1 = nullptr, *phi = nullptr, *tauPsi = nullptr;
  int *cFP = nullptr, *cEZ = nullptr, *nCF = nullptr, *nextC = nullptr, *numCorner = nullptr,
      *cOffSet = nullptr;
  int2* items = nullptr;
};

struct DropIn : DropBufs {
  hipStream_t stream = nullptr;
  size_t capC = 0, capZ = 0, capG = 0, capNb = 0, capCF = 0, capItems = 0;
};

constexpr int kMaxStreams = 80;
std::mutex g_dropMutex;
DropIn g_drop[kMaxStreams];

// Device scratch of the no-alloc drop-in (gpu_sweepucbxyz_noalloc): every
// device buffer of a call is carved from this array, whose address on each
// GPU is looked up once (hipGetSymbolAddress) with that GPU current.
constexpr size_t kNoAllocScratchBytes = 128ull << 20;
constexpr int kMaxDevices = 64;
void* g_scratchPtr[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// The current GPU's address of `scratch` (looked up on the first call on
// that GPU; callers hold g_dropMutex).
int scratch_address(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(53, "device %d beyond the no-alloc table", dev);
  if (!g_scratchPtr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(53, "no device address for the no-alloc scratch");
    g_scratchPtr[dev] = ptr;
  }
  *out = g_scratchPtr[dev];
  return 0;
}

// Bytes of the device buffers of a call, each rounded up to 256 bytes
// (mcf: corner faces per corner of the arrays; poly: the polyhedral path's
// extra arrays).
size_t dropin_bytes(size_t C, size_t Z, size_t G, size_t nb, bool tau, size_t mcf, bool poly) {
  auto r = [](size_t n) { return (n + 255) / 256 * 256; };
  return r(3 * mcf * C * 8) * 2 + r(C * 8) + r(G * C * 8) + r(G * Z * 8) + r(3 * 8) +
         r(G * (C + nb) * 8) + r(G * C * 8) + (tau ? r(G * C * 8) : 0) + r(mcf * C * 4) +
         r(Z * sizeof(int2)) + (poly ? r(mcf * C * 4) + 2 * r(C * 4) + 2 * r(Z * 4) : 0);
}

// Carves the buffers of a call from the no-alloc scratch.
int dropin_carve(DropBufs& b, size_t C, size_t Z, size_t G, size_t nb, bool tau, size_t mcf,
                 bool poly) {
  const size_t need = dropin_bytes(C, Z, G, nb, tau, mcf, poly);
  if (need > kNoAllocScratchBytes)
    return fail(53, "gpu_sweepucbxyz_noalloc: the call needs %zu bytes of device scratch, more "
                    "than its %zu (use gpu_sweepucbxyz)", need, kNoAllocScratchBytes);
  void* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  char* p = (char*)base;
  auto take = [&](auto*& dst, size_t bytes) {
    dst = (std::remove_reference_t<decltype(dst)>)(void*)p;
    p += (bytes + 255) / 256 * 256;
  };
  take(b.Afp, 3 * mcf * C * 8);
  take(b.Aez, 3 * mcf * C * 8);
  take(b.vol, C * 8);
  take(b.S, G * C * 8);
  take(b.sigt, G * Z * 8);
  take(b.omega, 3 * 8);
  take(b.psi1, G * (C + nb) * 8);
  take(b.phi, G * C * 8);
  if (tau) take(b.tauPsi, G * C * 8);
  else b.tauPsi = nullptr;
  take(b.cFP, mcf * C * 4);
  take(b.items, Z * sizeof(int2));
  if (poly) {
    take(b.cEZ, mcf * C * 4);
    take(b.nCF, C * 4);
    take(b.nextC, C * 4);
    take(b.numCorner, Z * 4);
    take(b.cOffSet, Z * 4);
  } else {
    b.cEZ = b.nCF = b.nextC = b.numCorner = b.cOffSet = nullptr;
  }
  return 0;
}

int dropin_alloc(DropIn& d, size_t C, size_t Z, size_t G, size_t nb, size_t mcf) {
  if (C <= d.capC && Z <= d.capZ && G <= d.capG && nb <= d.capNb && mcf <= d.capCF) return 0;
  dfree(d.Afp); dfree(d.Aez); dfree(d.vol); dfree(d.S); dfree(d.sigt); dfree(d.omega);
  dfree(d.psi1); dfree(d.phi); dfree(d.tauPsi); dfree(d.cFP); dfree(d.items);
  dfree(d.cEZ); dfree(d.nCF); dfree(d.nextC); dfree(d.numCorner); dfree(d.cOffSet);
  d.capC = std::max(C, d.capC); d.capZ = std::max(Z, d.capZ);
  d.capG = std::max(G, d.capG); d.capNb = std::max(nb, d.capNb);
  d.capCF = std::max(mcf, d.capCF);
  C = d.capC; Z = d.capZ; G = d.capG; nb = d.capNb; mcf = d.capCF;
  int rc;
  if ((rc = dalloc(&d.Afp, 3 * mcf * C)) || (rc = dalloc(&d.Aez, 3 * mcf * C)) ||
      (rc = dalloc(&d.vol, C)) || (rc = dalloc(&d.S, G * C)) || (rc = dalloc(&d.sigt, G * Z)) ||
      (rc = dalloc(&d.omega, 3)) || (rc = dalloc(&d.psi1, G * (C + nb))) ||
      (rc = dalloc(&d.phi, G * C)) || (rc = dalloc(&d.tauPsi, G * C)) ||
      (rc = dalloc(&d.cFP, mcf * C)) || (rc = dalloc(&d.items, Z)) ||
      (rc = dalloc(&d.cEZ, mcf * C)) || (rc = dalloc(&d.nCF, C)) || (rc = dalloc(&d.nextC, C)) ||
      (rc = dalloc(&d.numCorner, Z)) || (rc = dalloc(&d.cOffSet, Z))) {
    d.capC = d.capZ = d.capG = d.capNb = d.capCF = 0;
    return rc;
  }
  d.capItems = Z;
  return 0;
}

// Checks the corner arrays of a mesh that is not the canonical hex for the
// polyhedral kernel: every zone's corners inside (0, C], at most maxCorner of
// them, 3..maxcf corner faces per corner, EZ neighbours within the zone and
// nextC a permutation of the zone's corners. cornerZone receives each
// corner's zone.
int check_poly_topology(int Z, int C, int maxcf, int maxCorner, int nb, const int* numCorner,
                        const int* cOffSet, const int* nCFaces, const int* cEZ, const int* cFP,
                        const int* nextC, std::vector<int>& cornerZone) {
  cornerZone.assign(C, -1);
  for (int z = 0; z < Z; ++z) {
    const int n = numCorner[z], c0 = cOffSet[z];
    if (n < 1 || n > maxCorner || c0 < 0 || (long)c0 + n > C)
      return fail(11, "zone %d: %d corners at offset %d (maxCorner %d, %d corners)", z + 1, n, c0,
                  maxCorner, C);
    unsigned long long seen = 0;
    for (int c = 0; c < n; ++c) {
      const long cc = (long)c0 + c;
      if (cornerZone[cc] >= 0) return fail(11, "corner %ld belongs to two zones", cc + 1);
      cornerZone[cc] = z;
      if (nCFaces[cc] < 3 || nCFaces[cc] > maxcf)
        return fail(12, "corner %ld has %d corner faces (3 to maxcf = %d)", cc + 1, nCFaces[cc], maxcf);
      for (int f = 0; f < nCFaces[cc]; ++f) {
        const int e = cEZ[cc * maxcf + f];
        if (e < 1 || e > n || e == c + 1)
          return fail(13, "corner %ld face %d: EZ neighbour %d outside its zone's %d corners", cc + 1,
                      f + 1, e, n);
        if (cFP[cc * maxcf + f] < 1 || cFP[cc * maxcf + f] > C + nb)
          return fail(46, "cFP(%d, %ld) = %d out of range", f + 1, cc + 1, cFP[cc * maxcf + f]);
      }
      const int k = nextC[cc];
      if (k < 1 || k > n || (seen >> (k - 1) & 1))
        return fail(47, "zone %d: nextC is not an order of its %d corners", z + 1, n);
      seen |= 1ull << (k - 1);
    }
  }
  return 0;
}

// split_cyclic_planes for one angle of a polyhedral mesh: a zone goes one
// sub-level after the latest earlier zone of its plane that it shares a face
// with, so every read of a neighbour's psi sees what the serial sweep sees.
void split_cyclic_planes_poly(int Z, int C, int maxcf, const int* cFP, const int* nCFaces,
                              const int* numCorner, const int* cOffSet,
                              const std::vector<int>& cornerZone, std::vector<int2>& items,
                              std::vector<int>& planeOffset) {
  std::vector<int2> out;
  std::vector<int> offs(1, 0), mark(Z, -1), sub;
  const int L = (int)planeOffset.size() - 1;
  for (int h = 0; h < L; ++h) {
    const int i0 = planeOffset[h], i1 = planeOffset[h + 1];
    sub.assign(i1 - i0, 0);
    int maxSub = 0;
    for (int k = i0; k < i1; ++k) {
      const int z = items[k].x < 0 ? ~items[k].x : items[k].x;
      int sl = 0;
      for (int c = cOffSet[z]; c < cOffSet[z] + numCorner[z]; ++c)
        for (int f = 0; f < nCFaces[c]; ++f) {
          const int nb = cFP[(size_t)c * maxcf + f] - 1;
          if (nb >= C) continue;
          const int zn = cornerZone[nb];
          if (zn != z && zn >= 0 && mark[zn] >= 0) sl = std::max(sl, sub[mark[zn]] + 1);
        }
      sub[k - i0] = sl;
      mark[z] = k - i0;
      maxSub = std::max(maxSub, sl);
    }
    for (int sLevel = 0; sLevel <= maxSub; ++sLevel) {
      for (int k = i0; k < i1; ++k)
        if (sub[k - i0] == sLevel) out.push_back(items[k]);
      offs.push_back((int)out.size());
    }
    for (int k = i0; k < i1; ++k) mark[items[k].x < 0 ? ~items[k].x : items[k].x] = -1;
  }
  items.swap(out);
  planeOffset.swap(offs);
}

int dropin_sweep(int Angle, int nHP, const int* zip, const int* nextZ, const int* nextC,
                 const double* STotal, double tau, double* Psi, int G, const double* Volume,
                 const double* Sigt, const int* nCFaces, int ndim, int maxcf, int C,
                 const double* A_fp, const double* omega, const int* cFP, double* Psi1, int nb,
                 const double* A_ez, const int* cEZ, double quadwt, double* Phi, double* PsiB,
                 int maxCorner, int streamId, int savePsi, int numCycles, int cycleOffSet,
                 double* cyclePsi, const int* cycleList, int nBdyElem, int Mref,
                 const int* numCorner, const int* cOffSet, bool noAlloc) {
  if (ndim != 3) return fail(40, "only 3D (ndim = 3) is supported, got %d", ndim);
  if (maxcf < 3 || maxcf > umt::kPolyMaxCF || maxCorner < 1 || maxCorner > umt::kPolyMaxCorner)
    return fail(41, "maxcf = %d, maxCorner = %d (supported: maxcf 3 to %d, maxCorner up to %d)",
                maxcf, maxCorner, umt::kPolyMaxCF, umt::kPolyMaxCorner);
  if (numCycles < 0 || cycleOffSet < 0 || (numCycles > 0 && (!cyclePsi || !cycleList)))
    return fail(42, "angle %d: bad cycle list (numCycles %d, cycleOffSet %d)", Angle, numCycles, cycleOffSet);
  if (nBdyElem > 0 && Mref >= 1) return fail(43, "reflecting boundaries are not supported");
  if (nHP < 1 || G < 1 || C < 1 || nb < 0) return fail(44, "bad sizes");
  long Zl = 0;
  for (int h = 0; h < nHP; ++h) Zl += zip[h];
  const int Z = (int)Zl;
  // Canonical hex zones take the hex kernels; anything else the polyhedral
  // kernel (check_hex_topology's message is dropped when it refuses).
  const bool hex = maxcf == 3 && maxCorner == 8 && check_hex_topology(Z, C, numCorner, cOffSet, nCFaces, cEZ) == 0;
  std::vector<int> cornerZone;
  if (!hex) {
    if (!nextC) return fail(47, "the polyhedral sweep needs nextC");
    if (int rc = check_poly_topology(Z, C, maxcf, maxCorner, nb, numCorner, cOffSet, nCFaces, cEZ,
                                     cFP, nextC, cornerZone))
      return rc;
  }
  std::vector<int2> items(Z);
  for (int i = 0; i < Z; ++i) {
    const int z = nextZ[i];
    if (z == 0 || z > Z || z < -Z) return fail(45, "nextZ(%d) = %d out of range", i + 1, z);
    items[i] = make_int2(z > 0 ? z - 1 : ~(-z - 1), 0);  // cyclic zone: ~zone
  }
  if (hex)
    for (long i = 0; i < 3L * C; ++i)
      if (cFP[i] < 1 || cFP[i] > C + nb) return fail(46, "cFP entry %ld out of range", i + 1);
  for (int m = 0; m < numCycles; ++m)
    if (cycleList[cycleOffSet + m] < 1 || cycleList[cycleOffSet + m] > C)
      return fail(42, "angle %d: cycle list entry %d = corner %d out of range", Angle,
                  cycleOffSet + m + 1, cycleList[cycleOffSet + m]);
  std::vector<int> planeOffset(nHP + 1, 0);
  for (int h = 0; h < nHP; ++h) planeOffset[h + 1] = planeOffset[h] + zip[h];
  if (numCycles > 0) {
    if (hex)
      split_cyclic_planes(Z, C, cFP, 1, std::vector<int>(1, numCycles), items, planeOffset);
    else
      split_cyclic_planes_poly(Z, C, maxcf, cFP, nCFaces, numCorner, cOffSet, cornerZone, items,
                               planeOffset);
  }
  const int nLevels = (int)planeOffset.size() - 1;

  // Device buffers: the stream's own allocation (grown as needed), or for
  // the no-alloc entry point the static scratch on the null stream.
  const bool useTau = tau != 0.0;
  const size_t mcf = hex ? 3 : (size_t)maxcf;
  DropBufs d;
  hipStream_t s = nullptr;
  if (noAlloc) {
    if (int rc = dropin_carve(d, C, Z, G, nb, useTau, mcf, !hex)) return rc;
  } else {
    DropIn& di = g_drop[streamId % kMaxStreams];
    if (!di.stream) HIP_TRY(hipStreamCreate(&di.stream));
    if (int rc = dropin_alloc(di, C, Z, G, nb, mcf)) return rc;
    d = di;
    s = di.stream;
  }
  const size_t GC = (size_t)G * C;
  if (hex) {
    const std::vector<double> afpDev = umt::zone_blocked(A_fp, 9, (size_t)C);
    const std::vector<double> aezDev = umt::zone_blocked(A_ez, 9, (size_t)C);
    const std::vector<int> cfpDev = umt::zone_blocked(cFP, 3, (size_t)C);
    HIP_TRY(hipMemcpyAsync(d.Afp, afpDev.data(), 9 * (size_t)C * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.Aez, aezDev.data(), 9 * (size_t)C * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.cFP, cfpDev.data(), 3 * (size_t)C * 4, hipMemcpyHostToDevice, s));
  } else {
    HIP_TRY(hipMemcpyAsync(d.Afp, A_fp, 3 * mcf * C * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.Aez, A_ez, 3 * mcf * C * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.cFP, cFP, mcf * C * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.cEZ, cEZ, mcf * C * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.nCF, nCFaces, (size_t)C * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.nextC, nextC, (size_t)C * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.numCorner, numCorner, (size_t)Z * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d.cOffSet, cOffSet, (size_t)Z * 4, hipMemcpyHostToDevice, s));
  }
  HIP_TRY(hipMemcpyAsync(d.vol, Volume, (size_t)C * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d.S, STotal, GC * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d.sigt, Sigt, (size_t)G * Z * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d.omega, omega, 3 * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d.phi, Phi, GC * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d.items, items.data(), (size_t)Z * sizeof(int2), hipMemcpyHostToDevice, s));
  // The boundary tail of Psi1 starts from PsiB(:,:,Angle) (SweepUCBxyz.F90:95-97).
  if (nb > 0) std::memcpy(Psi1 + GC, PsiB, (size_t)G * nb * 8);
  // initFromCycleList (control/constructDynMemory.F90:118-152), on the host
  // copy of Psi1 that is uploaded whole.
  for (int m = cycleOffSet; m < cycleOffSet + numCycles; ++m)
    std::memcpy(Psi1 + (size_t)(cycleList[m] - 1) * G, cyclePsi + (size_t)m * G, (size_t)G * 8);
  HIP_TRY(hipMemcpyAsync(d.psi1, Psi1, (GC + (size_t)G * nb) * 8, hipMemcpyHostToDevice, s));
  if (useTau) HIP_TRY(hipMemcpyAsync(d.tauPsi, Psi, GC * 8, hipMemcpyHostToDevice, s));

  if (hex) {
    umt::SweepLevelArgs<double> a{};
    a.G = G;
    a.groupsPerLane = umt::groups_per_lane_supported(G, 2) ? 2 : 1;
    a.C = C;
    a.Afp = d.Afp;
    a.Aez = d.Aez;
    a.vol = d.vol;
    a.cFP = d.cFP;
    a.S = d.S;
    a.sigt = d.sigt;
    a.omega = d.omega;
    a.psi = d.psi1;
    a.psiAngleStride = 0;
    a.phi = d.phi;
    a.quadwt = quadwt;
    a.tauPsi = useTau ? d.tauPsi : nullptr;
    a.tau = tau;
    for (int h = 0; h < nLevels; ++h) {
      a.items = d.items + planeOffset[h];
      a.nItems = planeOffset[h + 1] - planeOffset[h];
      HIP_TRY(umt::launch_sweep_level<double>(a, true, s));
    }
  } else {
    umt::PolySweepArgs a{};
    a.G = G;
    a.C = C;
    a.maxcf = (int)mcf;
    a.numCorner = d.numCorner;
    a.cOffSet = d.cOffSet;
    a.nCFaces = d.nCF;
    a.cFP = d.cFP;
    a.cEZ = d.cEZ;
    a.nextC = d.nextC;
    a.Afp = d.Afp;
    a.Aez = d.Aez;
    a.vol = d.vol;
    a.S = d.S;
    a.sigt = d.sigt;
    a.omega = d.omega;
    a.tauPsi = useTau ? d.tauPsi : nullptr;
    a.tau = tau;
    a.psi = d.psi1;
    a.phi = d.phi;
    a.quadwt = quadwt;
    for (int h = 0; h < nLevels; ++h) {
      a.items = d.items + planeOffset[h];
      a.nItems = planeOffset[h + 1] - planeOffset[h];
      HIP_TRY(umt::launch_sweep_poly(a, s));
    }
  }
  HIP_TRY(hipMemcpyAsync(Psi1, d.psi1, GC * 8, hipMemcpyDeviceToHost, s));
  if (nb > 0) HIP_TRY(hipMemcpyAsync(PsiB, d.psi1 + GC, (size_t)G * nb * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(Phi, d.phi, GC * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (savePsi) std::memcpy(Psi, Psi1, GC * 8);
  // updateCycleList (control/constructDynMemory.F90:156-189)
  for (int m = cycleOffSet; m < cycleOffSet + numCycles; ++m)
    std::memcpy(cyclePsi + (size_t)m * G, Psi1 + (size_t)(cycleList[m] - 1) * G, (size_t)G * 8);
  return 0;
}

}  // namespace

void gpu_sweepucbxyz(int* Angle, int* nHyperPlanes, int* nZonesInPlane, int* nextZ, int* nextC,
                     double* STotal, double* tau, double* Psi, int* Groups, double* Volume,
                     double* Sigt, int* nCFacesArray, int* ndim, int* maxcf, int* ncorner,
                     double* A_fp, double* omega, int* cFP, double* Psi1, int* nbelem, double* A_ez,
                     int* cEZ, int* NumAngles, double* quadwt, double* Phi, double* PsiB,
                     int* maxCorner, int* mem0solve1, int* streamIdPtr, int* totalStreams,
                     int* savePsi, int* numCycles, int* cycleOffSet, double* cyclePsi,
                     int* cycleList, int* b0, int* nBdyElem, double* PsiBMref, int* Mref,
                     int* Geom_numCorner, int* Geom_cOffSet) {
  (void)NumAngles; (void)mem0solve1; (void)totalStreams; (void)b0; (void)PsiBMref;
  std::lock_guard<std::mutex> lock(g_dropMutex);
  g_dropin_status = dropin_sweep(*Angle, *nHyperPlanes, nZonesInPlane, nextZ, nextC, STotal, *tau, Psi,
                                 *Groups, Volume, Sigt, nCFacesArray, *ndim, *maxcf, *ncorner, A_fp,
                                 omega, cFP, Psi1, *nbelem, A_ez, cEZ, *quadwt, Phi, PsiB,
                                 *maxCorner, *streamIdPtr, *savePsi, *numCycles, *cycleOffSet,
                                 cyclePsi, cycleList, *nBdyElem, *Mref, Geom_numCorner,
                                 Geom_cOffSet, false);
  if (g_dropin_status) fprintf(stderr, "gpu_sweepucbxyz: %s\n", g_err.c_str());
}

void gpu_sweepucbxyz_noalloc(int* Angle, int* nHyperPlanes, int* nZonesInPlane, int* nextZ,
                             int* nextC, double* STotal, double* tau, double* Psi, int* Groups,
                             double* Volume, double* Sigt, int* nCFacesArray, int* ndim,
                             int* maxcf, int* ncorner, double* A_fp, double* omega, int* cFP,
                             double* Psi1, int* nbelem, double* A_ez, int* cEZ, int* NumAngles,
                             double* quadwt, double* Phi, double* PsiB, int* maxCorner,
                             int* mem0solve1, int* streamIdPtr, int* totalStreams, int* savePsi,
                             int* numCycles, int* cycleOffSet, double* cyclePsi, int* cycleList,
                             int* b0, int* nBdyElem, double* PsiBMref, int* Mref,
                             int* Geom_numCorner, int* Geom_cOffSet) {
  (void)NumAngles; (void)mem0solve1; (void)totalStreams; (void)b0; (void)PsiBMref;
  std::lock_guard<std::mutex> lock(g_dropMutex);
  g_dropin_status = dropin_sweep(*Angle, *nHyperPlanes, nZonesInPlane, nextZ, nextC, STotal, *tau, Psi,
                                 *Groups, Volume, Sigt, nCFacesArray, *ndim, *maxcf, *ncorner, A_fp,
                                 omega, cFP, Psi1, *nbelem, A_ez, cEZ, *quadwt, Phi, PsiB,
                                 *maxCorner, *streamIdPtr, *savePsi, *numCycles, *cycleOffSet,
                                 cyclePsi, cycleList, *nBdyElem, *Mref, Geom_numCorner,
                                 Geom_cOffSet, true);
  if (g_dropin_status) fprintf(stderr, "gpu_sweepucbxyz_noalloc: %s\n", g_err.c_str());
}

size_t gpu_sweepucbxyz_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

void gpu_streamsynchronize(int* streamId) {
  DropIn& d = g_drop[*streamId % kMaxStreams];
  if (d.stream) (void)hipStreamSynchronize(d.stream);
}

void gpu_devicesynchronize(void) { (void)hipDeviceSynchronize(); }

int gpu_sweepucbxyz_status(void) { return g_dropin_status; }

}  // extern "C"
