// Batches of uneven domains: many independent meshes (each with its own
// quadrature, group count and cross sections) swept together. Level h of a
// batched sweep holds level h of every domain (plane h of each of its angles
// in the current angle block), so one launch of ucb_sweep_batch_level covers
// all domains at that depth and the launch count is the deepest domain's
// level count, not the sum. A whole source iteration (source update, the
// level launches of every angle block with their phi reductions, and the
// residuals) is one hipGraph replay.
//
// Groups are stored padded to a multiple of 4 so that every groups-per-lane
// choice K in {1, 2, 4} divides them; padding groups have sigma_t = 1,
// sigma_s = q = 0, hence psi = phi = 0, and they are dropped on output and
// below the residual cutoff. The autotuner picks K and the angles per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/umt_sweep.h"
#include "common.h"
#include "host.h"
#include "kernels.h"

using umt::dalloc;
using umt::dfree;
using umt::fail;

// Problem access (runtime.hip owns the umt_problem layout).
namespace umt {
const Mesh& problem_mesh(const umt_problem* p);
const Geometry& problem_geometry(const umt_problem* p);
const Quadrature& problem_quadrature(const umt_problem* p);
const Schedule& problem_schedule(const umt_problem* p);
}  // namespace umt

namespace {

constexpr int kGroupPad = 4;
// Default bound on a workgroup's polls at a grid barrier of the persistent
// sweep: about a second, far above any level's time.
constexpr unsigned kBarrierPolls = 1u << 20;

struct Domain {
  int Z = 0, C = 0, nb = 0, G = 0, Gp = 0, A = 0;
  void* Afp = nullptr;
  void* Aez = nullptr;
  void* vol = nullptr;
  int* cFP = nullptr;
  double* vol64 = nullptr;
  double* volZ = nullptr;
  void* S = nullptr;     // (Gp, C)
  void* sigt = nullptr;  // (Gp, Z)
  double* sigs = nullptr;
  double* q = nullptr;
  double* phi = nullptr;  // (Gp, C)
  double* phibar = nullptr;
  double* phibarOld = nullptr;
  double* omega = nullptr;
  double* weight = nullptr;
  void* psi = nullptr;    // (Gp, C + nb, anglesPerBlock)
  int psiAngles = 0;
  bool haveXS = false;
  std::vector<int2> items;      // all-angle levels of this domain
  std::vector<int> levelOffset;
};

}  // namespace

struct umt_batch {
  int D = 0, precision = 64, K = 4, anglesPerBlock = 0, maxA = 0;
  hipStream_t stream = nullptr;
  bool useGraphs = true;
  std::vector<Domain> dom;
  void* descs = nullptr;             // BatchDomain<Real>[D]
  int2* items = nullptr;             // all blocks, all levels
  umt::BatchSeg* segs = nullptr;
  std::vector<long> levelItem;       // item offset per level (+ end)
  std::vector<int> levelSeg;         // segment offset per level (+ end)
  std::vector<long> levelTasks;
  std::vector<int> blockLevel;       // first level per angle block (+ end)
  unsigned long long* bits = nullptr;  // 2 per domain: max zone average, residual
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  long long unknowns = 0;
  // Persistent sweep: each angle block's levels in one launch of a grid the
  // device holds at once, with a grid barrier between levels.
  bool persistent = false;
  unsigned spinLimit = 0;            // polls at a barrier before a workgroup gives up
  int persistentBlocks = 0;          // workgroups of that launch (0: no resident grid)
  long* levelItemD = nullptr;        // levelItem, levelSeg, levelTasks on the device
  int* levelSegD = nullptr;
  long* levelTasksD = nullptr;
  umt::GridBarrier* bar = nullptr;
  size_t rsize() const { return precision == 32 ? sizeof(float) : sizeof(double); }
};

namespace {

void drop_graph(umt_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int upload_real(umt_batch* b, void* dst, const double* src, size_t n) {
  if (n == 0) return 0;
  if (b->precision == 64) {
    HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  } else {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
  }
  return 0;
}

// (G, n) -> (Gp, n) with `pad` in the padding groups.
std::vector<double> pad_groups(const double* a, int G, int Gp, size_t n, double pad) {
  std::vector<double> out((size_t)Gp * n, pad);
  for (size_t i = 0; i < n; ++i)
    for (int g = 0; g < G; ++g) out[i * Gp + g] = a[i * G + g];
  return out;
}

// All-angle hyperplane levels of one domain (plane h of every angle).
int domain_levels(const umt::Schedule& s, int Z, int A, Domain& d) {
  int maxPlanes = 0;
  for (int a = 0; a < A; ++a) maxPlanes = std::max(maxPlanes, s.nHyperPlanes[a]);
  std::vector<std::vector<int2>> lv(maxPlanes);
  for (int a = 0; a < A; ++a) {
    long done = 0;
    for (int h = 0; h < s.nHyperPlanes[a]; ++h) {
      const int n = s.zonesInPlane[s.planeOffset[a] + h];
      for (int i = 0; i < n; ++i) {
        const int z = s.nextZ[(long)a * Z + done + i];
        if (z <= 0 || z > Z) return fail(15, "batch: nextZ entry %d out of range", z);
        lv[h].push_back(make_int2(z - 1, a));
      }
      done += n;
    }
    if (done != Z) return fail(16, "batch: hyperplanes hold %ld zones, mesh has %d", done, Z);
  }
  d.items.clear();
  d.levelOffset.assign(1, 0);
  for (auto& l : lv) {
    // Zone order (then angle), as in the one-domain levels: a zone's
    // directions of one level read its source and geometry side by side.
    std::sort(l.begin(), l.end(),
              [](int2 p, int2 q) { return p.x != q.x ? p.x < q.x : p.y < q.y; });
    d.items.insert(d.items.end(), l.begin(), l.end());
    d.levelOffset.push_back((int)d.items.size());
  }
  return 0;
}

template <typename Real>
int upload_descs(umt_batch* b) {
  std::vector<umt::BatchDomain<Real>> h(b->D);
  for (int i = 0; i < b->D; ++i) {
    const Domain& d = b->dom[i];
    umt::BatchDomain<Real>& x = h[i];
    x.Afp = (const Real*)d.Afp;
    x.Aez = (const Real*)d.Aez;
    x.vol = (const Real*)d.vol;
    x.cFP = d.cFP;
    x.S = (const Real*)d.S;
    x.sigt = (const Real*)d.sigt;
    x.omega = d.omega;
    x.psi = (Real*)d.psi;
    x.psiAngleStride = (long)(d.C + d.nb) * d.Gp;
    x.G = d.Gp;
    x.C = d.C;
  }
  HIP_TRY(hipMemcpyAsync(b->descs, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

// Rebuilds the batched levels, segments and psi buffers for n angles per
// block and K groups per lane.
int layout(umt_batch* b, int n, int K) {
  n = std::max(1, std::min(n, b->maxA));
  HIP_TRY(hipStreamSynchronize(b->stream));
  drop_graph(b);
  std::vector<int2> items;
  std::vector<umt::BatchSeg> segs;
  std::vector<long> levelItem(1, 0), levelTasks;
  std::vector<int> levelSeg(1, 0), blockLevel(1, 0);
  const int nBlocks = (b->maxA + n - 1) / n;
  for (int blk = 0; blk < nBlocks; ++blk) {
    const int a0 = blk * n, a1 = a0 + n;
    int depth = 0;
    for (const Domain& d : b->dom) depth = std::max(depth, (int)d.levelOffset.size() - 1);
    for (int h = 0; h < depth; ++h) {
      const long itemBase = (long)items.size();
      long tasks = 0;
      for (int di = 0; di < b->D; ++di) {
        const Domain& d = b->dom[di];
        if (h >= (int)d.levelOffset.size() - 1 || a0 >= d.A) continue;
        const long before = (long)items.size();
        for (int i = d.levelOffset[h]; i < d.levelOffset[h + 1]; ++i) {
          const int2 it = d.items[i];
          if (it.y >= a0 && it.y < a1) items.push_back(it);
        }
        const long cnt = (long)items.size() - before;
        if (cnt == 0) continue;
        umt::BatchSeg sg;
        sg.taskStart = tasks;
        sg.domain = di;
        sg.itemStart = (int)(before - itemBase);
        sg.perItem = umt::make_fast_div((unsigned)(d.Gp / K));
        sg.nTasks = (int)(cnt * (d.Gp / K));
        segs.push_back(sg);
        // pad to whole waves so no wave spans two segments
        tasks += (sg.nTasks + umt::kBatchWaveTasks - 1) / umt::kBatchWaveTasks * umt::kBatchWaveTasks;
      }
      if (tasks == 0) continue;  // no domain has this level in this block
      levelItem.push_back((long)items.size());
      levelSeg.push_back((int)segs.size());
      levelTasks.push_back(tasks);
    }
    blockLevel.push_back((int)levelTasks.size());
  }
  dfree(b->items);
  dfree(b->segs);
  int rc;
  if ((rc = dalloc(&b->items, items.size())) || (rc = dalloc(&b->segs, segs.size()))) return rc;
  HIP_TRY(hipMemcpyAsync(b->items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->segs, segs.data(), segs.size() * sizeof(umt::BatchSeg), hipMemcpyHostToDevice, b->stream));
  for (Domain& d : b->dom) {
    const int na = std::min(n, d.A);
    if (na != d.psiAngles) {
      dfree(d.psi);
      const size_t cnt = (size_t)(d.C + d.nb) * d.Gp * na;
      if ((rc = dalloc((char**)&d.psi, cnt * b->rsize()))) return rc;
      if (b->precision == 32) HIP_TRY(umt::launch_fill_zero<float>((float*)d.psi, (long)cnt, b->stream));
      else HIP_TRY(umt::launch_fill_zero<double>((double*)d.psi, (long)cnt, b->stream));
      d.psiAngles = na;
    }
  }
  b->levelItem = levelItem;
  b->levelSeg = levelSeg;
  b->levelTasks = levelTasks;
  b->blockLevel = blockLevel;
  dfree(b->levelItemD);
  dfree(b->levelSegD);
  dfree(b->levelTasksD);
  if ((rc = dalloc(&b->levelItemD, levelItem.size())) || (rc = dalloc(&b->levelSegD, levelSeg.size())) ||
      (rc = dalloc(&b->levelTasksD, std::max<size_t>(1, levelTasks.size()))))
    return rc;
  HIP_TRY(hipMemcpyAsync(b->levelItemD, levelItem.data(), levelItem.size() * sizeof(long),
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->levelSegD, levelSeg.data(), levelSeg.size() * sizeof(int),
                         hipMemcpyHostToDevice, b->stream));
  if (!levelTasks.empty())
    HIP_TRY(hipMemcpyAsync(b->levelTasksD, levelTasks.data(), levelTasks.size() * sizeof(long),
                           hipMemcpyHostToDevice, b->stream));
  // The persistent grid: as many workgroups as the device holds at once,
  // and no more than the widest level has work for.
  constexpr long kWgTasks = umt::kPersistentThreads / 8;
  long widest = 1;
  for (long t : levelTasks) widest = std::max(widest, (t + kWgTasks - 1) / kWgTasks);
  const int resident = umt::batch_persistent_resident_blocks(b->precision, K);
  b->persistentBlocks = (int)std::min<long>(resident, widest);
  b->anglesPerBlock = n;
  b->K = K;
  if ((rc = b->precision == 32 ? upload_descs<float>(b) : upload_descs<double>(b))) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

// Launches of one source iteration on the batch's stream: S from phi, the
// sweep (levels and phi per angle block), the residual per domain.
template <typename Real>
int launch_iteration(umt_batch* b, bool withResidual) {
  for (Domain& d : b->dom)
    HIP_TRY(umt::launch_source_update<Real>(d.phi, d.sigs, d.q, d.Gp, (long)d.Gp * d.C, (Real*)d.S,
                                            b->stream));
  for (int blk = 0; blk + 1 < (int)b->blockLevel.size(); ++blk) {
    const int a0 = blk * b->anglesPerBlock;
    if (b->persistent) {
      umt::BatchPersistentArgs<Real> a{};
      a.doms = (const umt::BatchDomain<Real>*)b->descs;
      a.items = b->items;
      a.segs = b->segs;
      a.levelItem = b->levelItemD;
      a.levelSeg = b->levelSegD;
      a.levelTasks = b->levelTasksD;
      a.h0 = b->blockLevel[blk];
      a.h1 = b->blockLevel[blk + 1];
      a.angle0 = a0;
      a.bar = b->bar;
      a.spinLimit = b->spinLimit;
      if (a.h1 > a.h0)
        HIP_TRY(umt::launch_sweep_batch_persistent<Real>(a, b->K, b->persistentBlocks, b->stream));
    } else {
      for (int h = b->blockLevel[blk]; h < b->blockLevel[blk + 1]; ++h) {
        umt::BatchLevelArgs<Real> a{};
        a.doms = (const umt::BatchDomain<Real>*)b->descs;
        a.items = b->items + b->levelItem[h];
        a.segs = b->segs + b->levelSeg[h];
        a.nSegs = b->levelSeg[h + 1] - b->levelSeg[h];
        a.nTasks = b->levelTasks[h];
        a.angle0 = a0;
        HIP_TRY(umt::launch_sweep_batch_level<Real>(a, b->K, b->stream));
      }
    }
    for (Domain& d : b->dom) {
      if (a0 >= d.A) continue;
      const int nA = std::min(b->anglesPerBlock, d.A - a0);
      HIP_TRY(umt::launch_phi_from_psi_acc<Real>((const Real*)d.psi, d.weight + a0, nA,
                                                 (long)(d.C + d.nb) * d.Gp, (long)d.Gp * d.C,
                                                 d.phi, blk > 0, b->stream));
    }
  }
  if (withResidual)
    for (int i = 0; i < b->D; ++i) {
      Domain& d = b->dom[i];
      HIP_TRY(umt::launch_residual(d.phi, d.vol64, d.volZ, d.Gp, (long)d.Gp * d.Z, d.phibar,
                                   d.phibarOld, 1e-6, b->bits + 2 * i, b->stream));
    }
  return 0;
}

int iteration(umt_batch* b, bool withResidual) {
  for (const Domain& d : b->dom)
    if (!d.haveXS) return fail(50, "batch: every domain needs umt_batch_set_cross_sections first");
  if (!b->useGraphs || !withResidual)
    return b->precision == 32 ? launch_iteration<float>(b, withResidual)
                              : launch_iteration<double>(b, withResidual);
  if (!b->exec) {
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = b->precision == 32 ? launch_iteration<float>(b, true) : launch_iteration<double>(b, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_TRY(e);
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  return 0;
}

// After a persistent sweep: fails (and clears the barrier for the next
// launch) if a workgroup gave up waiting at a grid barrier.
int barrier_status(umt_batch* b) {
  if (!b->persistent) return 0;
  umt::GridBarrier h{};
  HIP_TRY(hipMemcpyAsync(&h, b->bar, sizeof h, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (!h.err) return 0;
  HIP_TRY(hipMemsetAsync(b->bar, 0, sizeof(umt::GridBarrier), b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return fail(60, "batch: a workgroup of the persistent sweep gave up after %u polls at a grid "
                  "barrier (the grid was not resident at once); this sweep's results are invalid: "
                  "use per-level launches (umt_batch_set_persistent(b, 0, 0))", b->spinLimit);
}

void free_domain(Domain& d) {
  dfree(d.Afp); dfree(d.Aez); dfree(d.vol); dfree(d.cFP); dfree(d.vol64); dfree(d.volZ);
  dfree(d.S); dfree(d.sigt); dfree(d.sigs); dfree(d.q); dfree(d.phi); dfree(d.phibar);
  dfree(d.phibarOld); dfree(d.omega); dfree(d.weight); dfree(d.psi);
}

int create_domain(umt_batch* b, const umt_problem* p, int G, Domain& d) {
  const umt::Mesh& m = umt::problem_mesh(p);
  const umt::Geometry& g = umt::problem_geometry(p);
  const umt::Quadrature& q = umt::problem_quadrature(p);
  const umt::Schedule& s = umt::problem_schedule(p);
  if (int rc = umt::check_hex_topology(m.nzones, m.ncorner, m.numCorner.data(), m.cOffSet.data(),
                                       m.nCFaces.data(), m.cEZ.data()))
    return rc;
  d.Z = m.nzones; d.C = m.ncorner; d.nb = m.nbelem; d.G = G; d.A = q.nangles;
  d.Gp = (G + kGroupPad - 1) / kGroupPad * kGroupPad;
  if (s.totalCycles() > 0)
    return fail(51, "batch: a domain's schedule has dependency cycles (%ld cycle-list corners); "
                    "sweep it with its own context", s.totalCycles());
  if (int rc = domain_levels(s, d.Z, d.A, d)) return rc;
  const size_t rs = b->rsize(), C = d.C, Z = d.Z, Gp = d.Gp;
  int rc;
  if ((rc = dalloc((char**)&d.Afp, 9 * C * rs)) || (rc = dalloc((char**)&d.Aez, 9 * C * rs)) ||
      (rc = dalloc((char**)&d.vol, C * rs)) || (rc = dalloc(&d.cFP, 3 * C)) ||
      (rc = dalloc(&d.vol64, C)) || (rc = dalloc(&d.volZ, Z)) ||
      (rc = dalloc((char**)&d.S, Gp * C * rs)) || (rc = dalloc((char**)&d.sigt, Gp * Z * rs)) ||
      (rc = dalloc(&d.sigs, Gp * Z)) || (rc = dalloc(&d.q, Gp * Z)) || (rc = dalloc(&d.phi, Gp * C)) ||
      (rc = dalloc(&d.phibar, Gp * Z)) || (rc = dalloc(&d.phibarOld, Gp * Z)) ||
      (rc = dalloc(&d.omega, 3 * (size_t)d.A)) || (rc = dalloc(&d.weight, (size_t)d.A)))
    return rc;
  if ((rc = upload_real(b, d.Afp, umt::zone_blocked(g.A_fp.data(), 9, C).data(), 9 * C)) ||
      (rc = upload_real(b, d.Aez, umt::zone_blocked(g.A_ez.data(), 9, C).data(), 9 * C)) ||
      (rc = upload_real(b, d.vol, g.Volume.data(), C)))
    return rc;
  const std::vector<int> cfpDev = umt::zone_blocked(m.cFP.data(), 3, C);
  HIP_TRY(hipMemcpyAsync(d.cFP, cfpDev.data(), 3 * C * sizeof(int), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipMemcpyAsync(d.vol64, g.Volume.data(), C * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(d.volZ, g.VolumeZone.data(), Z * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(d.omega, q.omega.data(), 3 * (size_t)d.A * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(d.weight, q.weight.data(), (size_t)d.A * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(umt::launch_fill_zero<double>(d.phi, (long)(Gp * C), b->stream));
  HIP_TRY(umt::launch_fill_zero<double>(d.phibarOld, (long)(Gp * Z), b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->unknowns += (long long)d.C * d.A * d.G;
  return 0;
}

}  // namespace

extern "C" {

void umt_batch_destroy(umt_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (Domain& d : b->dom) free_domain(d);
  dfree(b->descs); dfree(b->items); dfree(b->segs); dfree(b->bits);
  dfree(b->levelItemD); dfree(b->levelSegD); dfree(b->levelTasksD); dfree(b->bar);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int umt_batch_create(int nDomains, const umt_problem* const* problems, const int* groups,
                     int precision, umt_batch** out) {
  if (!out || !problems || !groups || nDomains < 1) return fail(1, "bad argument");
  if (precision != 32 && precision != 64) return fail(1, "precision must be 32 or 64");
  *out = nullptr;
  umt_batch* b = new umt_batch();
  auto bail = [&](int rc) {
    std::string keep = umt_last_error();
    umt_batch_destroy(b);
    fail(rc, "%s", keep.c_str());
    return rc;
  };
  b->D = nDomains;
  b->precision = precision;
  b->dom.resize(nDomains);
  if (hipStreamCreate(&b->stream) != hipSuccess) return bail(fail(2, "hipStreamCreate failed"));
  if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(2, "hipEventCreate failed"));
  for (int i = 0; i < nDomains; ++i) {
    if (!problems[i] || groups[i] < 1) return bail(fail(1, "domain %d: bad problem or group count", i));
    if (int rc = create_domain(b, problems[i], groups[i], b->dom[i])) return bail(rc);
    b->maxA = std::max(b->maxA, b->dom[i].A);
  }
  int rc;
  if ((rc = dalloc((char**)&b->descs, (size_t)nDomains * (precision == 32 ? sizeof(umt::BatchDomain<float>)
                                                                          : sizeof(umt::BatchDomain<double>)))) ||
      (rc = dalloc(&b->bits, 2 * (size_t)nDomains)) || (rc = dalloc(&b->bar, 1)))
    return bail(rc);
  if (hipMemsetAsync(b->bar, 0, sizeof(umt::GridBarrier), b->stream) != hipSuccess)
    return bail(fail(2, "hipMemsetAsync failed"));
  b->spinLimit = kBarrierPolls;
  // Start with blocks of at most 32 angles (psi for all angles of every
  // domain may not fit); the autotuner tries larger blocks where they fit.
  if ((rc = layout(b, std::min(b->maxA, 32), 4))) return bail(rc);
  *out = b;
  return 0;
}

int umt_batch_dims(const umt_batch* b, int64_t* d) {
  if (!b || !d) return fail(1, "null argument");
  d[0] = b->D;
  d[1] = b->unknowns;
  d[2] = b->persistent ? (int64_t)b->blockLevel.size() - 1 : (int64_t)b->levelTasks.size();
  d[3] = b->K;
  d[4] = b->anglesPerBlock;
  d[5] = (int64_t)b->blockLevel.size() - 1;
  return 0;
}

int umt_batch_set_cross_sections(umt_batch* b, int domain, const double* sigt, const double* sigs,
                                 const double* q) {
  if (!b || !sigt || !sigs || !q) return fail(1, "null argument");
  if (domain < 0 || domain >= b->D) return fail(1, "domain %d out of range", domain);
  Domain& d = b->dom[domain];
  const size_t n = (size_t)d.G * d.Z;
  for (size_t i = 0; i < n; ++i)
    if (!(sigt[i] > 0.0) || !std::isfinite(sigt[i]) || !(sigs[i] >= 0.0) || !std::isfinite(sigs[i]) ||
        !std::isfinite(q[i]))
      return fail(30, "domain %d: cross-section entry %zu is not valid", domain, i + 1);
  for (size_t i = 0; i < n; ++i)
    if (sigs[i] > sigt[i])
      return fail(52, "domain %d: sigs entry %zu exceeds Sigt (scattering ratio above 1)", domain, i + 1);
  const std::vector<double> st = pad_groups(sigt, d.G, d.Gp, d.Z, 1.0);
  const std::vector<double> ss = pad_groups(sigs, d.G, d.Gp, d.Z, 0.0);
  const std::vector<double> qq = pad_groups(q, d.G, d.Gp, d.Z, 0.0);
  if (int rc = upload_real(b, d.sigt, st.data(), st.size())) return rc;
  HIP_TRY(hipMemcpyAsync(d.sigs, ss.data(), ss.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(d.q, qq.data(), qq.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  d.haveXS = true;
  return 0;
}

int umt_batch_reset(umt_batch* b) {
  if (!b) return fail(1, "null batch");
  for (Domain& d : b->dom) {
    HIP_TRY(umt::launch_fill_zero<double>(d.phi, (long)d.Gp * d.C, b->stream));
    HIP_TRY(umt::launch_fill_zero<double>(d.phibarOld, (long)d.Gp * d.Z, b->stream));
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

int umt_batch_get_phi(const umt_batch* b, int domain, double* phi) {
  if (!b || !phi) return fail(1, "null argument");
  if (domain < 0 || domain >= b->D) return fail(1, "domain %d out of range", domain);
  const Domain& d = b->dom[domain];
  std::vector<double> h((size_t)d.Gp * d.C);
  HIP_TRY(hipMemcpyAsync(h.data(), d.phi, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t c = 0; c < (size_t)d.C; ++c)
    for (int g = 0; g < d.G; ++g) phi[c * d.G + g] = h[c * d.Gp + g];
  return 0;
}

int umt_batch_sweep(umt_batch* b) {
  if (!b) return fail(1, "null batch");
  if (int rc = iteration(b, false)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  return barrier_status(b);
}

int umt_batch_set_persistent(umt_batch* b, int on, int maxPolls) {
  if (!b) return fail(1, "null batch");
  if (on && b->persistentBlocks < 1)
    return fail(61, "batch: the device reports no resident grid for the persistent sweep");
  HIP_TRY(hipStreamSynchronize(b->stream));
  drop_graph(b);
  b->persistent = on != 0;
  b->spinLimit = maxPolls > 0 ? (unsigned)maxPolls : maxPolls < 0 ? 0u : kBarrierPolls;
  return 0;
}

int umt_batch_get_persistent(const umt_batch* b, int* on, int* blocks) {
  if (!b) return fail(1, "null batch");
  if (on) *on = b->persistent ? 1 : 0;
  if (blocks) *blocks = b->persistentBlocks;
  return 0;
}

int umt_batch_set_graphs(umt_batch* b, int on) {
  if (!b) return fail(1, "null batch");
  b->useGraphs = on != 0;
  return 0;
}

int umt_batch_set_layout(umt_batch* b, int groupsPerLane, int anglesPerBlock) {
  if (!b) return fail(1, "null batch");
  if (groupsPerLane != 1 && groupsPerLane != 2 && groupsPerLane != 4)
    return fail(1, "groups per lane must be 1, 2 or 4");
  if (anglesPerBlock < 1) return fail(1, "angles per block must be at least 1");
  return layout(b, anglesPerBlock, groupsPerLane);
}

int umt_batch_source_iteration(umt_batch* b, double tol, int maxIters, int* iters,
                               double* residuals, int* totalIters) {
  if (!b) return fail(1, "null batch");
  std::vector<int> done(b->D, 0);
  std::vector<double> res(b->D, 1.0);
  std::vector<unsigned long long> bits(2 * (size_t)b->D);
  int it = 0;
  bool all = false;
  // Zone averages of the starting phi are the "old" values of iteration 1.
  for (int i = 0; i < b->D; ++i) {
    Domain& d = b->dom[i];
    HIP_TRY(umt::launch_residual(d.phi, d.vol64, d.volZ, d.Gp, (long)d.Gp * d.Z, d.phibar,
                                 d.phibarOld, 1e-6, b->bits + 2 * i, b->stream));
  }
  while (it < maxIters && !all) {
    if (int rc = iteration(b, true)) return rc;
    HIP_TRY(hipMemcpyAsync(bits.data(), b->bits, bits.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (int rc = barrier_status(b)) return rc;
    ++it;
    all = true;
    for (int i = 0; i < b->D; ++i) {
      double r;
      std::memcpy(&r, &bits[2 * i + 1], sizeof r);
      if (!std::isfinite(r)) return fail(34, "batch: domain %d residual is not finite at iteration %d", i, it);
      res[i] = r;
      if (!done[i] && r < tol) done[i] = it;
      all = all && done[i];
    }
  }
  for (int i = 0; i < b->D; ++i) {
    if (iters) iters[i] = done[i] ? done[i] : it;
    if (residuals) residuals[i] = res[i];
  }
  if (totalIters) *totalIters = it;
  return all ? 0 : 1;
}

int umt_batch_time_iterations(umt_batch* b, int warmup, int reps, double* msPerIteration) {
  if (!b || reps < 1) return fail(1, "bad argument");
  for (int i = 0; i < warmup; ++i)
    if (int rc = iteration(b, true)) return rc;
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  for (int i = 0; i < reps; ++i)
    if (int rc = iteration(b, true)) return rc;
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  HIP_TRY(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (msPerIteration) *msPerIteration = ms / reps;
  return barrier_status(b);
}

int umt_batch_autotune(umt_batch* b, int reps, int* bestK, int* bestAnglesPerBlock, double* msBest) {
  if (!b || reps < 1) return fail(1, "bad argument");
  const int Ks[3] = {1, 2, 4};
  std::vector<int> blocks;
  for (int n : {8, 16, 32, 64, b->maxA})
    if (n <= b->maxA && std::find(blocks.begin(), blocks.end(), n) == blocks.end()) blocks.push_back(n);
  double best = HUGE_VAL;
  int bk = b->K, bn = b->anglesPerBlock;
  if (int rc = umt_batch_set_persistent(b, 0, 0)) return rc;  // candidates with per-level launches
  for (int n : blocks) {
    // Skip block sizes whose psi buffers would not fit in free memory.
    size_t need = 0;
    for (const Domain& d : b->dom) need += (size_t)(d.C + d.nb) * d.Gp * std::min(n, d.A) * b->rsize();
    size_t freeB = 0, totalB = 0;
    HIP_TRY(hipMemGetInfo(&freeB, &totalB));
    size_t held = 0;
    for (const Domain& d : b->dom) held += (size_t)(d.C + d.nb) * d.Gp * d.psiAngles * b->rsize();
    if (need > freeB + held) continue;
    for (int K : Ks) {
      if (int rc = layout(b, n, K)) return rc;
      double ms = 0.0;
      if (int rc = umt_batch_time_iterations(b, 1, reps, &ms)) return rc;
      if (ms < best) {
        best = ms;
        bk = K;
        bn = n;
      }
    }
  }
  if (int rc = layout(b, bn, bk)) return rc;
  // The persistent sweep at that layout, kept if faster.
  if (b->persistentBlocks > 0) {
    if (int rc = umt_batch_set_persistent(b, 1, 0)) return rc;
    double ms = 0.0;
    if (int rc = umt_batch_time_iterations(b, 1, reps, &ms)) return rc;
    if (ms < best) best = ms;
    else if (int rc = umt_batch_set_persistent(b, 0, 0)) return rc;
  }
  if (int rc = umt_batch_reset(b)) return rc;
  if (bestK) *bestK = bk;
  if (bestAnglesPerBlock) *bestAnglesPerBlock = bn;
  if (msBest) *msBest = best;
  return 0;
}

}  // extern "C"
