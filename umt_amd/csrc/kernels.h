// Kernel argument blocks and launch wrappers for the UCB sweep kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace umt {

// Quotient of n < 2^31 by an invariant divisor d >= 1 as one multiply and a
// shift: q = (n m) >> s with s = 31 + ceil(log2 d) and m = floor(2^s / d) + 1,
// exact because 2^s < m d <= 2^s + 2^(s - 31) (Granlund and Montgomery,
// "Division by invariant integers using multiplication", 1994, Thm 4.2).
// Kernels use it to split a flat thread index into (item, group block).
struct FastDiv {
  unsigned long long m;
  int shift;
};
inline FastDiv make_fast_div(unsigned d) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  FastDiv f;
  f.shift = 31 + l;
  f.m = (1ull << f.shift) / d + 1;
  return f;
}
__device__ __forceinline__ unsigned fast_div(unsigned n, FastDiv f) {
  return (unsigned)(((unsigned long long)n * f.m) >> f.shift);
}
// Level launches index threads with 32 bits.
constexpr long kMaxLevelThreads = 0x7fffffffL;

template <typename Real>
struct SweepLevelArgs {
  const int2* items;        // (zone, angle) pairs of this level, 0-based; a zone
                            // with cyclic corner dependencies comes as ~zone
  int nItems;
  int G;                    // groups
  int groupsPerLane;        // K: groups one lane carries (1, 2, 4 or 8; divides G)
  int C;                    // corners; cFP values above C are boundary elements
  const Real* Afp;          // zone-blocked: A_fp(d, f, 8 z + c) at (9 z + 3 f + d) 8 + c
  const Real* Aez;          // zone-blocked like Afp
  const Real* vol;          // (C)
  const int* cFP;           // zone-blocked: cFP(f, 8 z + c) at (3 z + f) 8 + c, 1-based
  const Real* S;            // (G, C) source per steradian
  const Real* sigt;         // (G, Z)
  const double* omega;      // (3, A)
  Real* psi;                // (G, C + nbelem, A) all angles, or one angle
  long psiAngleStride;      // (C + nbelem) * G
  int angle0;               // first angle held in psi (angle blocks)
  Real parTol;              // corner faces with |Omega.A| <= parTol |A|_1 are parallel (0: exact)
  int noInflow;             // every zone on its own: incident FP fluxes are zero
  int boundaryIn;           // all-angle sweeps: incoming boundary faces read psi's
                            // boundary tail (incident flux), else they bring zero
  // Single-angle (drop-in) extras.
  double* phi;              // (G, C), phi += quadwt * psi
  double quadwt;
  const Real* tauPsi;       // null, or source += tau * Psi: (G, C) of this angle (single
                            // angle) or the previous cycle's (G, C + nbelem, A) (all angles)
  Real tau;
  FastDiv perItem;          // division by G / groupsPerLane (set by launch_sweep_level)
};

template <typename Real>
hipError_t launch_sweep_level(const SweepLevelArgs<Real>& a, bool singleAngle, hipStream_t s);

// One hyperplane of the drop-in's single-angle FP64 sweep on a mesh of
// general polyhedral zones (corner arrays in Teton's layout, not
// zone-blocked). One thread per (zone, group).
constexpr int kPolyMaxCorner = 32;  // corners per zone
constexpr int kPolyMaxCF = 8;       // corner faces per corner
struct PolySweepArgs {
  const int2* items;        // (zone, 0) of this hyperplane, 0-based; a zone on the
  int nItems;               // cycle list comes as ~zone
  int G;
  int C;                    // corners; cFP values above C are boundary elements
  int maxcf;                // leading dimension of the corner-face arrays
  const int* numCorner;     // (Z)
  const int* cOffSet;       // (Z)
  const int* nCFaces;       // (C)
  const int* cFP;           // (maxcf, C), 1-based
  const int* cEZ;           // (maxcf, C), 1-based corner within the zone
  const int* nextC;         // (C): corner order of each zone, 1-based within the zone
  const double* Afp;        // (3, maxcf, C)
  const double* Aez;        // (3, maxcf, C)
  const double* vol;        // (C)
  const double* S;          // (G, C)
  const double* sigt;       // (G, Z)
  const double* omega;      // (3)
  const double* tauPsi;     // null, or source += tau * Psi (G, C)
  double tau;
  double* psi;              // (G, C + nbelem): corner fluxes, then the boundary tail
  double* phi;              // (G, C), phi += quadwt * psi
  double quadwt;
};
hipError_t launch_sweep_poly(const PolySweepArgs& a, hipStream_t s);

// Many domains in one level launch (batches of uneven domains).
template <typename Real>
struct BatchDomain {
  const Real* Afp;          // zone-blocked (as SweepLevelArgs) of this domain
  const Real* Aez;
  const Real* vol;
  const int* cFP;           // zone-blocked, 1-based within the domain
  const Real* S;            // (G, C)
  const Real* sigt;         // (G, Z)
  const double* omega;      // (3, A)
  Real* psi;                // (G, C + nbelem, anglesPerBlock)
  long psiAngleStride;      // (C + nbelem) * G
  int G;                    // groups as stored (a multiple of the groups per lane)
  int C;
};
// Tasks of one wave (64 lanes, 8 per task) stay in one segment: segments
// start at multiples of kBatchWaveTasks and are padded to them.
constexpr int kBatchWaveTasks = 8;
struct BatchSeg {
  long taskStart;           // first task (item, K groups) of the segment, a multiple of 8
  int domain;
  int itemStart;            // first item of the segment within the level
  FastDiv perItem;          // division by the domain's G / K
  int nTasks;               // tasks of the segment before padding
};
template <typename Real>
struct BatchLevelArgs {
  const BatchDomain<Real>* doms;
  const int2* items;        // (zone, angle) of this level, grouped by domain
  const BatchSeg* segs;     // segments of this level
  int nSegs;
  long nTasks;              // padded tasks of the level
  int angle0;               // first angle of the block being swept
};
template <typename Real>
hipError_t launch_sweep_batch_level(const BatchLevelArgs<Real>& b, int K, hipStream_t s);
// Barrier state of the persistent batch kernel (device memory, zeroed before
// first use; count returns to 0 after every completed barrier).
struct GridBarrier {
  unsigned count;           // groups of workgroups arrived at the current barrier
  unsigned gen;             // barriers completed
  unsigned err;             // nonzero: a workgroup gave up waiting (grid not resident)
  unsigned pad[29];
  unsigned sub[8][32];      // workgroups w with w mod 8 = i arrived: sub[i][0] (one cache line each)
};
// Threads of a workgroup of the persistent batch kernel (one workgroup per
// CU at its occupancy, so that few workgroups meet at each grid barrier).
constexpr int kPersistentThreads = 1024;
// All levels [h0, h1) of one angle block of a batch in one launch.
template <typename Real>
struct BatchPersistentArgs {
  const BatchDomain<Real>* doms;
  const int2* items;        // every level's items
  const BatchSeg* segs;     // every level's segments
  const long* levelItem;    // first item of level h
  const int* levelSeg;      // first segment of level h (h + 1: end)
  const long* levelTasks;   // padded tasks of level h
  int h0, h1;
  int angle0;
  GridBarrier* bar;
  unsigned spinLimit;       // polls a workgroup waits at a barrier before giving up
};
template <typename Real>
hipError_t launch_sweep_batch_persistent(const BatchPersistentArgs<Real>& b, int K, int blocks,
                                         hipStream_t s);
// Workgroups of the persistent batch kernel the device holds at once.
int batch_persistent_resident_blocks(int precision, int K);
// Nonzero when the sweep kernels can carry K groups per lane for G groups.
int groups_per_lane_supported(int G, int K);
// Layout code of the xyz zone layout (ucb_sweep_level_zone): all eight
// corners of an item on one lane, one group per lane, W = 64, 32 or 16
// lanes (the largest dividing G) per item. The corner layout's codes are its
// groups per lane, K = 1, 2, 4, 8.
constexpr int kZoneLayout = -1;
// Lanes per item of the zone layout for G groups (0: G is not a multiple of 16).
int zone_lanes_for(int G);
// Nonzero when an xyz all-angle sweep can run layout code K for G groups.
int sweep_layout_supported(int G, int K);
template <typename Real>
hipError_t launch_phi_from_psi(const Real* psi, const double* w, int A, long stride, long n,
                               double* phi, hipStream_t s);
template <typename Real>
hipError_t launch_source_update(const double* phi, const double* sigs, const double* q, int G,
                                long n, Real* S, hipStream_t s);
hipError_t launch_residual(const double* phi, const double* vol, const double* volZ, int G, long nzg,
                           double* phibar, double* phibarOld, double cutoff,
                           unsigned long long* bits, hipStream_t s, bool magnitude = false);
template <typename T>
hipError_t launch_fill_zero(T* p, long n, hipStream_t s);
hipError_t launch_copy_probe(const double* src, double* dst, long n, hipStream_t s);
// phi = sum_a w_a psi (accumulate: phi += ...).
template <typename Real>
hipError_t launch_phi_from_psi_acc(const Real* psi, const double* w, int A, long stride, long n,
                                   double* phi, bool accumulate, hipStream_t s);

// Grey-transport acceleration (gta_kernels.hip).
hipError_t launch_gta_grey_source(const double* phiHalf, const double* phiPrev, const double* sigs,
                                  int G, long n, double* gsrc, hipStream_t s);
template <typename Real>
hipError_t launch_gta_sweep_source(int kind, const double* P, const double* other,
                                   const double* sigs, const double* eps, int G, long n, Real* S,
                                   hipStream_t s);
hipError_t launch_gta_lincomb(const double* x, const double* y, const double* z,
                              const double* coef, int G, long n, double* out, hipStream_t s);
// Transfer-matrix GTA solver: unit sources, per-zone LU factors of
// I - wtiso sigma_s T, and the operator P + (LU)^-1 (phi - P).
template <typename Real>
hipError_t launch_gta_unit_source(int s, int G, long n, Real* S, hipStream_t st);
hipError_t launch_gta_zone_lu(double* T, const double* sigs, int G, long nzg, long CG,
                              hipStream_t s);
// residual: write P - M(P) instead of M(P).
hipError_t launch_gta_zone_solve(const double* LU, const double* P, const double* phi, int G,
                                 long nzg, long CG, double* out, bool residual, hipStream_t s);
int gta_dot_block(int G);
hipError_t launch_gta_wdot(const double* x, const double* y, const double* sigs, const double* vol,
                           int G, long n, int nblocks, double* partial, double* out, hipStream_t s);
// out[0..G) = sum x y sigma_s V, out[G..2G) = sum x x sigma_s V (partial: 2 nblocks G).
hipError_t launch_gta_wdot2(const double* x, const double* y, const double* sigs, const double* vol,
                            int G, long n, int nblocks, double* partial, double* out, hipStream_t s);
// F += alpha d + omega r; r -= omega As (coef[3g], coef[3g+1]); out = sum r sigma_s V.
hipError_t launch_gta_bicg_update(double* F, const double* d, double* r, const double* As,
                                  const double* coef, const double* sigs, const double* vol, int G,
                                  long n, int nblocks, double* partial, double* out, hipStream_t s);
hipError_t launch_gta_zone_error(const double* F, const double* phiHalf, const double* vol,
                                 const double* volZ, int G, long nzg, double* pzOld, int nblocks,
                                 double* partial, double* out, hipStream_t s);
template <typename Dst, typename Src>
hipError_t launch_convert(const Src* s, Dst* d, long n, hipStream_t s_);
// Cycle-list psi: entries (corner, angle) first .. first + n - 1; toPsi loads
// cyclePsi (G, entries) into psi (G, C + nbelem, angles from angle0), else
// stores psi into cyclePsi.
// Incident boundary fluxes of angles a0 .. a0+nA-1 into the boundary tails
// of their psi blocks: psi[(a - a0) angleStride + C G + i] = inc[a nb G + i].
template <typename Real>
hipError_t launch_boundary_in(const Real* inc, long nbG, long CG, int a0, int nA, long angleStride,
                              Real* psi, hipStream_t s);
template <typename Real>
hipError_t launch_cycle_copy(const int2* entries, long first, long n, int G, Real* cyclePsi,
                             Real* psi, long angleStride, int angle0, bool toPsi, hipStream_t s);
// Zone averages phibar(g,z) = sum_c V_c phi(g,c) / V_z (maxBits: scratch running max).
hipError_t launch_zone_average(const double* phi, const double* vol, const double* volZ, int G,
                               long nzg, double* phibar, unsigned long long* maxBits,
                               hipStream_t s);

// 2D cylindrical (RZ) sweep (rz_kernels.hip).
struct RzAngle {
  double mu, xi;        // direction cosines along r and z
  double fac;           // angDerivFac
  double w1, w2;        // quadTauW1, quadTauW2: PsiM <- w1 psi - w2 PsiM
  int level;            // xi level (index of its half-angle flux buffer)
  int flags;            // 1: starting direction, 2: the next direction is finishing
};
template <typename Real>
struct RzLevelArgs {
  const int2* items;    // (zone, direction) pairs of this level, 0-based
  int nItems;
  int G;                // groups (a multiple of the groups per lane)
  int C;                // corners; cFP values above C are boundary elements
  const Real* Afp;      // (2, 2, C)
  const Real* Aez;      // (2, 2, C)
  const Real* Rfp;      // (2, C) RadiusFP
  const Real* Rez;      // (2, C) RadiusEZ
  const Real* area;     // (C)
  const Real* vol;      // (C) r-weighted corner volume
  const int* cFP;       // (2, C), 1-based
  const Real* S;        // (G, C)
  const Real* sigt;     // (G, Z)
  const RzAngle* angles;
  Real* psi;            // (G, C + nbelem, A)
  long psiAngleStride;  // (C + nbelem) * G
  Real* psiM;           // (G, C, xi levels)
  const Real* tauPsi;   // null, or the previous cycle's psi (G, C, A): source += tau * it
  Real tau;
  FastDiv perItem;      // division by G / K (set by launch_rz_sweep_level)
};
template <typename Real>
hipError_t launch_rz_sweep_level(const RzLevelArgs<Real>& a, int K, hipStream_t s);
// Tile levels: one 64-lane wave per (tile, direction) item and chunk of
// gbPerChunk (1, 2 or 4) group blocks; items/order/stepOff/nsteps are this
// level's.
template <typename Real>
struct RzTileArgs {
  RzLevelArgs<Real> p;            // the per-corner arrays (items unused)
  const int2* items;              // (tile, direction) of this level
  const unsigned char* order;     // (items, maxZones) zone slots sorted by step
  const unsigned char* stepOff;   // (items, maxZones + 1) first entry of each step
  const int* nsteps;              // (items)
  const int* tileZone;            // (tiles, maxZones) zone or -1
  int maxZones;
  int gbPerChunk;
};
template <typename Real>
hipError_t launch_rz_sweep_tile(const RzTileArgs<Real>& a, int nItems, int K, hipStream_t s);
template <typename Real>
hipError_t launch_rz_source_update(const double* phi, const double* sigs, const double* q, int G,
                                   long n, Real* S, hipStream_t s);
hipError_t launch_rz_residual(const double* phi, const double* vol, const double* volZ, int G,
                              long nzg, double* phibar, double* phibarOld, double cutoff,
                              unsigned long long* bits, hipStream_t s);

// Learned correction network (learn_kernels.hip): 16 feature slots -> 64 ->
// 64 -> 1 with tanh. FP32 parameter layout: W1 [feature][hidden], b1,
// W2 [in][out], b2, w3, b3.
constexpr int kMlpFeatures = 16;
constexpr int kMlpHidden = 64;
constexpr int kMlpUsedFeatures = 10;
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpFeatures * kMlpHidden;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;
constexpr int kMlpB3 = kMlpW3 + kMlpHidden;
constexpr int kMlpParams = kMlpB3 + 1;
constexpr int kMlpPacked = 20 * 64 * 8;  // BF16 operand slots (mlp_pack)
constexpr double kLearnTiny = 1e-300;
constexpr int kLearnTargetBlocks = 256;

hipError_t launch_mlp_pack(const float* params, unsigned short* packed, hipStream_t s);
hipError_t launch_mlp_forward(const float* X, long n, const unsigned short* packed,
                              const float* params, float* r, hipStream_t s);
// Waves (a multiple of 4) mlp_train uses for count samples; partial needs
// that many rows of kMlpParams + 1 floats.
int mlp_train_waves(long count, int stepsPerWave);
hipError_t launch_mlp_train(const float* X, const float* a, const float* b, const int* idx,
                            long start, long count, int stepsPerWave,
                            const unsigned short* packed, const float* params, float* partial,
                            hipStream_t s);
hipError_t launch_mlp_reduce_adam(const float* partial, int nWaves, float* grad, float* params,
                                  float* m, float* v, float lr, float beta1, float beta2, float eps,
                                  float bc1, float bc2, hipStream_t s);
hipError_t launch_learn_features(const double* ph, const double* pp, const double* dp,
                                 const double* sigt, const double* sigs, const double* q,
                                 const int* zoneNbr, int G, long n, double* rel, double* ratio,
                                 unsigned long long* maxRel, float* X, hipStream_t s);
hipError_t launch_learn_apply(double* phi, const double* phiPrev, const float* r, int G, long n,
                              unsigned long long* bad, hipStream_t s);
hipError_t launch_learn_delta(const double* ph, const double* pp, long n, double* d, hipStream_t s);
hipError_t launch_learn_targets(const double* ph, const double* d, const double* pstar, long n,
                                double* a, double* b, double* partial, float* aOut, float* bOut,
                                hipStream_t s);

}  // namespace umt
