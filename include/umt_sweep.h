/*
 * C-ABI of the MI355X UCB xyz sweep library (libumt_amd.so).
 *
 * Every entry point takes plain pointers and sizes. Arrays use the Teton
 * layout: column-major shapes (first index fastest, groups fastest), index
 * contents 1-based (corner, zone, face, boundary-element numbers).
 * Functions returning int return 0 on success and a nonzero code otherwise;
 * umt_last_error() gives the message of the calling thread's last failure.
 * Nothing here aborts or exits the process.
 */
#ifndef UMT_SWEEP_H
#define UMT_SWEEP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Drop-in sweep, one angle of one phase-space set per call.
 *
 * Replaces the CUDA entry point of the same name and argument list,
 * src/teton/gpu/GPU_SweepUCBxyz.cu:532-572, bound from Fortran with
 * bind(c) in src/teton/gpu/SweepUCBxyzToGPU.F90:43-91 and called at
 * SweepUCBxyzToGPU.F90:198-240. Semantics follow snac/SweepUCBxyz.F90:11-322:
 * the boundary tail of Psi1 is loaded from PsiB(:,:,Angle), Psi1 receives the
 * corner fluxes, Phi += quadwt * Psi1, exiting boundary fluxes go to PsiB, and
 * Psi(:,:,Angle) = Psi1 when *savePsi != 0. Outputs are complete when the call
 * returns (gpu_streamsynchronize is kept for the caller's sequence).
 * Supported: 3D meshes of any polyhedral zones, up to 32 corners per zone
 * and 8 corner faces per corner (*maxCorner <= 32, 3 <= *maxcf <= 8). Hex
 * zones in the corner-face order setTetonZone produces (*maxcf 3,
 * *maxCorner 8) take the hex-tuned kernels; any other mesh takes a general
 * kernel that follows nextC within each zone and reproduces the serial
 * FP64 sweep bit for bit. Acyclic schedules and schedules with cycle lists
 * (*numCycles corners of cycleList from *cycleOffSet load Psi1 from cyclePsi
 * before the sweep and store back after it,
 * control/constructDynMemory.F90:118-189; cyclic zones come as negative
 * nextZ entries), no reflecting boundary (*nBdyElem == 0 or *Mref < 1).
 * Anything else leaves the outputs untouched and sets a nonzero
 * gpu_sweepucbxyz_status().
 * ---------------------------------------------------------------------- */
void gpu_sweepucbxyz(int *Angle, int *nHyperPlanes, int *nZonesInPlane, int *nextZ, int *nextC,
                     double *STotal, double *tau, double *Psi, int *Groups, double *Volume,
                     double *Sigt, int *nCFacesArray, int *ndim, int *maxcf, int *ncorner,
                     double *A_fp, double *omega, int *cFP, double *Psi1, int *nbelem,
                     double *A_ez, int *cEZ, int *NumAngles, double *quadwt, double *Phi,
                     double *PsiB, int *maxCorner, int *mem0solve1, int *streamIdPtr,
                     int *totalStreams, int *savePsi, int *numCycles, int *cycleOffSet,
                     double *cyclePsi, int *cycleList, int *b0, int *nBdyElem, double *PsiBMref,
                     int *Mref, int *Geom_numCorner, int *Geom_cOffSet);

/* The same sweep as gpu_sweepucbxyz, same arguments and results, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in
 * the reference. It allocates nothing: every device buffer of the call is
 * carved from one static device array of gpu_sweepucbxyz_noalloc_max_bytes()
 * bytes (128 MiB) that each GPU holds until the process ends; calls whose
 * buffers would not fit are refused (status 53). It runs on the null stream
 * and its outputs are complete when it returns. Every call on one GPU shares
 * that array, which keeps the last call's contents (each call clears what it
 * accumulates into). Calls that overlapped on one GPU would race on the
 * array: the library's drop-in lock serialises them within a process (each
 * process holds its own copy of the array). */
void gpu_sweepucbxyz_noalloc(int *Angle, int *nHyperPlanes, int *nZonesInPlane, int *nextZ,
                             int *nextC, double *STotal, double *tau, double *Psi, int *Groups,
                             double *Volume, double *Sigt, int *nCFacesArray, int *ndim,
                             int *maxcf, int *ncorner, double *A_fp, double *omega, int *cFP,
                             double *Psi1, int *nbelem, double *A_ez, int *cEZ, int *NumAngles,
                             double *quadwt, double *Phi, double *PsiB, int *maxCorner,
                             int *mem0solve1, int *streamIdPtr, int *totalStreams, int *savePsi,
                             int *numCycles, int *cycleOffSet, double *cyclePsi, int *cycleList,
                             int *b0, int *nBdyElem, double *PsiBMref, int *Mref,
                             int *Geom_numCorner, int *Geom_cOffSet);
/* Bytes of the no-alloc drop-in's device scratch. */
size_t gpu_sweepucbxyz_noalloc_max_bytes(void);

/* Replaces gpu_streamsynchronize, GPU_SweepUCBxyz.cu:520-523. */
void gpu_streamsynchronize(int *streamId);

/* Replaces gpu_devicesynchronize, GPU_SweepUCBxyz.cu:525-528. */
void gpu_devicesynchronize(void);

/* Status of the calling thread's last gpu_sweepucbxyz call (0 = success).
 * No counterpart in the reference, whose wrapper exits the process on error
 * (GPU_SweepUCBxyz.cu:18-27); the argument list above is kept unchanged. */
int gpu_sweepucbxyz_status(void);

/* Message of the calling thread's last failure, or "" if none. */
const char *umt_last_error(void);

/* ------------------------------------------------------------------------
 * Problem construction on the host (mesh, geometry, quadrature, schedule).
 * ---------------------------------------------------------------------- */
typedef struct umt_problem umt_problem;

/* Tiled hex mesh of [0,1]^3 (nx*ny*nz tiles of 24 zones, UMT's tile,
 * driver/test_driver.cc:1787-1816), corner connectivity as
 * aux/setTetonZone.F90:143-191, corner geometry as rt/geometryUCBxyz.F90,
 * product quadrature (rt/quadProduct.F90, polar axis x) with
 * npolar x nazimuthal directions per octant, and the per-angle schedule of
 * rt/rtorder.F90 / snac/snnext.F90. kdtree != 0 orders zones by a k-d tree. */
int umt_problem_tiled(int nx, int ny, int nz, int npolar, int nazimuthal, int kdtree,
                      umt_problem **out);
/* A problem on any hex mesh given as explicit points coords(3, npoints) and
 * 8 node ids per zone, connectivity(8, nzones), 0-based, in the Conduit
 * Blueprint / VTK hexahedron order (nodes 1-4 one face, counter-clockwise
 * seen from nodes 5-8 above them): the unstructured-topology form UMT hands
 * to Teton (interface/TetonBlueprint.cc:774-890), turned into Teton's corner
 * arrays as aux/setTetonZone.F90:143-191 does. kdtree != 0 reorders zones
 * as the UMT driver does. Meshes whose S_N schedule has dependency cycles
 * are refused; where only the S2 grey-acceleration schedule has cycles (S2
 * directions lying in warped faces) the problem is built without GTA and
 * umt_problem_gta says why. */
int umt_problem_from_hex_mesh(int npoints, const double *coords, int nzones,
                              const int *connectivity, int npolar, int nazimuthal, int kdtree,
                              umt_problem **out);
void umt_problem_destroy(umt_problem *p);

/* dims[0..7] = nzones, ncorner, nbelem, nangles, maxPlanes, npoints,
 *              total planes over all angles, maxcf (= 3). */
int umt_problem_dims(const umt_problem *p, int64_t *dims);

/* Mesh arrays (Geometry_mod names): numCorner(Z), cOffSet(Z), nCFacesArray(C),
 * cFP(3,C), cEZ(3,C), CToFace(3,C), zoneOpp(6,Z), faceOpp(6,Z), px(3,C).
 * Any output pointer may be NULL. */
int umt_problem_mesh(const umt_problem *p, int *numCorner, int *cOffSet, int *nCFaces, int *cFP,
                     int *cEZ, int *CToFace, int *zoneOpp, int *faceOpp, double *px);

/* A_fp(3,3,C), A_ez(3,3,C), Volume(C), VolumeZone(Z). */
int umt_problem_geometry(const umt_problem *p, double *A_fp, double *A_ez, double *Volume,
                         double *VolumeZone);

/* omega(3,A), weight(A) with sum(weight) = 4 pi. */
int umt_problem_quadrature(const umt_problem *p, double *omega, double *weight);

/* nHyperPlanes(A), zonesInPlane (all angles concatenated), nextZ(Z,A),
 * nextC(C,A) (AngleSet_mod names). */
int umt_problem_schedule(const umt_problem *p, int *nHyperPlanes, int *zonesInPlane, int *nextZ,
                         int *nextC);

/* The S_N schedule's cycle lists (ASet%numCycles, ASet%cycleList of
 * mods/AngleSet_mod.F90:80-82, built as rt/rtorder.F90:62-90 builds them):
 * *totalCycles, numCycles(A), cycleList(totalCycles) with 1-based corners,
 * angle by angle. Any output may be NULL (query totalCycles first). Meshes
 * whose schedule has no dependency cycles have totalCycles == 0; zones with
 * cyclic corner dependencies appear negated in nextZ
 * (snac/snnext.F90:164-168). */
int umt_problem_cycles(const umt_problem *p, int *totalCycles, int *numCycles, int *cycleList);

/* Grey-acceleration quadrature (level-symmetric S2, rt/quadxyz.F90 order 2,
 * the GTA order of driver/test_driver.cc:1206) and its schedule. Faces that
 * contain an S2 direction (the tile's diagonal faces) carry no dependency
 * for it. nangles receives 8; array sizes as umt_problem_schedule. */
int umt_problem_gta(const umt_problem *p, int *nangles, double *omega, double *weight,
                    int *nHyperPlanes, int *zonesInPlane, int *nextZ);

/* Seeded cross sections per (zone, group), counter-based so any ordering of
 * the draws gives the same values: sigt ~ log-uniform[sigt_lo, sigt_hi],
 * sigs = c * sigt with c ~ U[c_lo, c_hi], q ~ U[q_lo, q_hi]. Arrays (G, Z). */
int umt_draw_cross_sections(uint64_t seed, int nzones, int groups, double sigt_lo,
                            double sigt_hi, double c_lo, double c_hi, double q_lo, double q_hi,
                            double *sigt, double *sigs, double *q);

/* ------------------------------------------------------------------------
 * Device-resident sweep context: geometry, schedule (hyperplane zone
 * batches over all angles), cross sections, psi for all angles and phi stay
 * in HBM. precision = 64 (FP64 sweep) or 32 (FP32 sweep; phi, the
 * source-iteration residual and balance stay FP64).
 * ---------------------------------------------------------------------- */
typedef struct umt_ctx umt_ctx;

/* From Teton-layout arrays (as a Fortran host holds them). zonesInPlane is
 * the concatenation over angles of each angle's nHyperPlanes entries.
 * nextZ may hold cyclic zones (negative entries, snac/snnext.F90:164-168);
 * sweeps then need the schedule's cycle lists (umt_ctx_set_cycles). */
int umt_ctx_create(int nzones, int ncorner, int nbelem, int groups, int nangles, int precision,
                   const int *numCorner, const int *cOffSet, const int *nCFaces, const int *cFP,
                   const int *cEZ, const double *A_fp, const double *A_ez, const double *Volume,
                   const double *VolumeZone, const double *omega, const double *weight,
                   const int *nHyperPlanes, const int *zonesInPlane, const int *nextZ,
                   umt_ctx **out);
int umt_ctx_create_from_problem(const umt_problem *p, int groups, int precision, umt_ctx **out);
void umt_ctx_destroy(umt_ctx *ctx);

/* The schedule's cycle lists: numCycles(A) and cycleList (1-based corners,
 * angle by angle; ASet%numCycles / ASet%cycleList, mods/AngleSet_mod.F90:80-82,
 * rt/rtorder.F90:62-90). Each sweep loads the listed corners' psi from the
 * previous sweep of their angle before the angle is swept and stores them
 * after it, as SetSweep.F90:149,155 do with initFromCycleList /
 * updateCycleList (control/constructDynMemory.F90:118-189); a cyclic zone
 * is solved once from its own corners' previous values
 * (SweepUCBxyz.F90:281-295). The lagged values start at zero;
 * umt_ctx_set_psi_prev reloads them from the previous cycle's Psi
 * (initCyclePsi, control/constructDynMemory.F90:75-114) and
 * umt_ctx_copy_phi carries them between contexts with the same lists.
 * umt_ctx_create_from_problem sets them from the problem's schedule. */
int umt_ctx_set_cycles(umt_ctx *ctx, const int *numCycles, const int *cycleList);

/* Sigt(G,Z) = total cross section seen by the sweep (setTotalOpacity with tau = 0). */
int umt_ctx_set_sigt(umt_ctx *ctx, const double *sigt);
/* STotal(G,C): source per steradian for the next sweep. */
int umt_ctx_set_source(umt_ctx *ctx, const double *STotal);
/* Scattering cross section sigs(G,Z) and fixed source q(G,Z) for source iteration. */
int umt_ctx_set_scattering(umt_ctx *ctx, const double *sigs, const double *q);

/* One sweep of all angles with the current source: psi for every angle and
 * phi = sum_a w_a psi. Boundaries are vacuum (incoming psi_B = 0,
 * control/setBoundarySources.F90:76). */
int umt_ctx_sweep(umt_ctx *ctx);
int umt_ctx_get_phi(const umt_ctx *ctx, double *phi);   /* (G,C) */
/* psi(G, C+nbelem, A): corner fluxes, then the boundary tail (exiting fluxes
 * for exiting boundary elements, zero for incoming ones). */
int umt_ctx_get_psi(const umt_ctx *ctx, double *psi);
int umt_ctx_set_phi(umt_ctx *ctx, const double *phi);   /* (G,C) */

/* Source iteration: S = (sigs phi + q)/(4 pi); sweep; phi from psi; residual =
 * max over (zone, group) of |phibar_new - phibar_old| / phibar_new for
 * phibar_new > 1e-6 * max(phibar_new) (rt/ConvergenceTest.F90:150-168), zone
 * averages volume-weighted. Starts from the phi held by the context. Stops
 * when residual < tol or after maxIters iterations; residual_history (may be
 * NULL) receives up to maxIters values. Returns 0 when converged, 1 when
 * maxIters was reached (the state is still valid), 2 when the stall window
 * ended it, 34 when the residual is not finite (a NaN or Inf zone average:
 * the iterates diverged) and 52 when some sigs exceeds Sigt. */
int umt_ctx_source_iteration(umt_ctx *ctx, double tol, int maxIters, int *iters,
                             double *residual, double *residual_history);
/* Time-dependent sweeps. tau = 1/(c dt) (control/initializeSets.F90:81):
 * the sweep sees Sigt + tau (control/setTotalOpacity.F90) and the source
 * STotal + tau Psi(g, c, Angle) (snac/SweepUCBxyz.F90:121), Psi being the
 * previous cycle's angular flux held on the device for every angle. tau = 0
 * (default) is the steady problem. umt_ctx_save_psi is the cycle's final
 * sweep with savePsi (rt/rtmainsn.F90:238-246, SweepUCBxyz.F90:314-318):
 * source from the current phi, one sweep, Psi = its psi, phi = its phi.
 * Psi arrays on the host are Teton's Psi(G, C, A). GTA accounts for tau in
 * its grey total cross section. */
/* Incident boundary fluxes for the context's S_N sweeps: PsiB(G, nbelem, A)
 * as Teton's PsiB, loaded into each angle's boundary tail before its levels
 * (snac/SweepUCBxyz.F90:95-97, control/setBoundarySources.F90); NULL returns
 * to vacuum boundaries (the default). umt_ctx_get_boundary_flux gives
 * PsiB(G, nbelem, A) after a sweep: exiting fluxes in outgoing slots,
 * incident ones in the others (needs one angle block). */
int umt_ctx_set_incident_flux(umt_ctx *ctx, const double *PsiB);
int umt_ctx_get_boundary_flux(const umt_ctx *ctx, double *PsiB);
int umt_ctx_set_tau(umt_ctx *ctx, double tau);
int umt_ctx_set_psi_prev(umt_ctx *ctx, const double *Psi);
int umt_ctx_get_psi_prev(const umt_ctx *ctx, double *Psi);
int umt_ctx_save_psi(umt_ctx *ctx);

/* window > 0: source iteration also stops (returning 2) once `window`
 * iterations pass without a new smallest residual, as at the rounding floor
 * of FP32 sweeps; 0 (default) never stops early. No counterpart in the reference. */
int umt_ctx_set_stall_window(umt_ctx *ctx, int window);

/* Timing: warmup + reps sweeps of the current source, timed with HIP events
 * on the context's stream. ms_sweep = mean wall time of one full sweep
 * (level kernels + phi reduction); ms_level_kernels = mean summed duration
 * of the sweep-level launches in one sweep; launches = level launches per sweep. */
int umt_ctx_time_sweeps(umt_ctx *ctx, int warmup, int reps, double *ms_sweep,
                        double *ms_level_kernels, int *launches);

/* Grey-transport acceleration of source iteration (rt/GTASolver.F90): per
 * group, the scattering-rate change after each sweep drives a BiCGStab solve
 * of the S2 grey-correction problem, each step one S2 sweep plus one
 * stretched-TSA sweep (rt/GreySweep.F90), and the correction is added to
 * phi. umt_ctx_enable_gta takes the S2 directions and their schedule
 * (umt_problem_gta gives them; umt_ctx_create_from_problem enables GTA
 * itself). The S2 directions lie in mesh faces: the schedule must treat
 * corner faces with |Omega.A| <= 1e-12 |A|_1 as parallel, as the S2 sweeps
 * do (1e-5 in FP32 sweeps), which umt_problem_gta's schedule does. Replaces the GTA path of rt/LinearSolver.F90:87-121, which the
 * mini-app build compiles out. */
int umt_ctx_enable_gta(umt_ctx *ctx, int nangles, const double *omega, const double *weight,
                       const int *nHyperPlanes, const int *zonesInPlane, const int *nextZ);
/* mode 0: plain source iteration (default); 1: GTA-accelerated; 2: learned
 * correction (umt_ctx_set_learner); 3: both, the learned correction applied to
 * the GTA-corrected iterates (GTA is then the fallback and the arbiter). */
int umt_ctx_set_acceleration(umt_ctx *ctx, int mode);
/* Autotuner of the S2 grey sweeps' layout (their levels are much smaller
 * than the S_N sweep's, so the best layout can differ): times groups per
 * lane K = 1, 2, 4, 8 and UMT_ZONE_LAYOUT (msPerLayout[0..4], -1 when
 * unsupported), keeps the fastest. Needs the cross sections. No
 * counterpart in the reference. */
int umt_ctx_autotune_grey_layouts(umt_ctx *ctx, int reps, int *bestK, double *msPerLayout);
/* The same, reporting the four corner-layout timings only (msPerK[0..3]);
 * *bestK can still be UMT_ZONE_LAYOUT. */
int umt_ctx_autotune_grey(umt_ctx *ctx, int reps, int *bestK, double *msPerK);
/* Cumulative BiCGStab iterations and S2 sweeps done by GTA on this context. */
int umt_ctx_gta_stats(const umt_ctx *ctx, int64_t *greyIterations, int64_t *greySweeps);
/* The grey operator on host arrays (groups, corners): out = M(P), the S2
 * grey sweep of wtiso (sigma_s P + gsrc) followed by the stretched-TSA sweep
 * (rt/GreySweep.F90:60-96); P or gsrc NULL stand for zero. */
int umt_ctx_grey_sweep(umt_ctx *ctx, const double *P, const double *gsrc, double *out);
/* One grey solve (rt/GTASolver.F90:185-402) for the grey source gsrc after a
 * sweep that gave phiHalf: F = the grey correction; greyIterations = the
 * BiCGStab iterations it took. */
int umt_ctx_grey_solve(umt_ctx *ctx, const double *gsrc, const double *phiHalf, double epsGrey,
                       double *F, int64_t *greyIterations);
/* GTA solver kernel, numbered as the reference driver's option
 * (driver/test_driver.cc:1520-1531): 1 = grey S2 sweep followed by a
 * stretched-TSA sweep per operator application (rt/GreySweep.F90:50-110,
 * the default, as in the reference); 2 = the transfer-matrix solver
 * (Size%useNewGTASolver): per-zone corner transfer matrices
 * (snac/InitSweepGreyUCBxyz.F90:10-253) and each zone's direct solve of its
 * grey scattering problem with in-place LU factors
 * (snac/UpdateScalarIntensity.F90:11-170) after one grey sweep
 * (rt/GreySweep.F90:12-47), the factors made once per set of opacities. */
int umt_ctx_set_gta_solver(umt_ctx *ctx, int kernel);
/* Grey iteration controls: relative tolerance (grey_reltol, default 1e-6)
 * and maximum grey iterations (grey_max_sweeps, default 21). Replace
 * teton_adjust_grey_reltol and teton_adjust_grey_maxits
 * (aux/AdjustGreyRelTol.F90:11, aux/AdjustGreyMaxIts.F90:11). Status 57
 * for a tolerance <= 0 or fewer than one iteration. */
int umt_ctx_set_grey_controls(umt_ctx *ctx, double relTol, int maxIts);
int umt_ctx_get_grey_controls(const umt_ctx *ctx, double *relTol, int *maxIts,
                              int *enforceHardIterMax);
/* enforceHardIterMax != 0: the grey iteration stops at its maximum whatever
 * its error (rt/GTASolver.F90:373-375). Replaces the first argument of
 * teton_setgtaoptions (aux/setGTAOptions.F90:15-34); the reference's default,
 * kept here, is off for kernel 1 and on for kernel 2
 * (mods/GreyAcceleration_mod.F90:122,169), set again by umt_ctx_set_gta_solver.
 * forceExtraOuter, the second argument, is not built: it needs the grey
 * control's convergence state, which GTASolver.F90 never sets. */
int umt_ctx_set_gta_options(umt_ctx *ctx, int enforceHardIterMax);
/* Kernel 2's matrices for checking: T(g, c, s) (Fortran order, G x C x 8)
 * is the scalar flux at corner c (1..C) from a unit per-steradian source at
 * local corner s (1..8) of the same zone, summed over the grey directions
 * (GTA%TT(s, c), one per group); LU (same shape, may be NULL) holds, for
 * element (row c, column s) of each zone's matrix I - wtiso sigma_s T, its
 * Doolittle factors (unit lower below the diagonal, upper on and above). */
int umt_ctx_grey_transfer(umt_ctx *ctx, double *T, double *LU);

/* Layout of the all-angle sweep kernels (default 1). K = 1, 2, 4 or 8
 * (dividing the group count): one lane per corner of a zone, carrying K
 * groups. K = UMT_ZONE_LAYOUT (-1, group count a multiple of 16): one lane
 * per group of a zone, carrying its eight corners. The per-angle drop-in
 * always runs the first. No counterpart in the reference. */
#define UMT_ZONE_LAYOUT (-1)
int umt_ctx_set_groups_per_lane(umt_ctx *ctx, int K);
int umt_ctx_get_groups_per_lane(const umt_ctx *ctx, int *K);

/* on != 0: replay each sweep's level launches from a captured hipGraph
 * (captured on first use per source, cross section and groups per lane; the
 * graph holds kernel nodes only). No counterpart in the reference. */
int umt_ctx_set_graphs(umt_ctx *ctx, int on);

/* Sweep the angles in blocks of n (default: all in one block). Each block
 * runs its own hyperplane levels and reuses one psi buffer of n angles, so
 * psi memory is (C + nbelem) * G * n values; phi sums the blocks in angle
 * order. With more than one block umt_ctx_get_psi is unavailable. No
 * counterpart in the reference (its angle sets, mods/QuadratureList_mod.F90,
 * play the same role for threads). */
int umt_ctx_set_angles_per_block(umt_ctx *ctx, int n);
int umt_ctx_get_angles_per_block(const umt_ctx *ctx, int *n, int *blocks);

/* dst's phi = src's phi, device to device (same groups and corners), and
 * the lagged cycle-list psi when both have the same cycle lists. */
int umt_ctx_copy_phi(umt_ctx *dst, const umt_ctx *src);

/* Autotuner: times reps sweeps with each supported layout (groups per lane
 * 1, 2, 4, 8 and the zone layout) on the context's own problem and keeps
 * the fastest. msPerK[0..3] receives the
 * mean ms per sweep for K = 1, 2, 4, 8 (-1 when unsupported); may be NULL.
 * No counterpart in the reference. */
int umt_ctx_autotune(umt_ctx *ctx, int reps, int *bestK, double *msPerK);
/* The same, reporting msPerLayout[0..4] for K = 1, 2, 4, 8 and
 * UMT_ZONE_LAYOUT (-1 when unsupported). No counterpart in the reference. */
int umt_ctx_autotune_layouts(umt_ctx *ctx, int reps, int *bestK, double *msPerLayout);

/* n sweeps back to back on the context's stream, one synchronisation at the end. */
int umt_ctx_sweep_n(umt_ctx *ctx, int n);

/* Mean incoming interior FP corner faces per (corner, angle): the n_in of
 * the algorithmic-bytes model (Omega.A_fp < 0 towards a corner, not a
 * boundary element), counted over all corners and angles. */
int umt_ctx_incoming_faces(const umt_ctx *ctx, double *n_in);

/* Level structure of the S_N sweep as launched: stats[0] level launches per
 * sweep, stats[1] (zone, angle) items per sweep, stats[2] distinct (level,
 * zone) pairs (a zone met by several directions in one level launch counts
 * once: those directions share its source, cross section and geometry
 * reads), stats[3] angle blocks. No counterpart in the reference. */
int umt_ctx_level_stats(const umt_ctx *ctx, int64_t *stats);

/* Free and total device memory of the current GPU in bytes (hipMemGetInfo),
 * for sizing angle blocks. No counterpart in the reference. */
int umt_device_memory_info(int64_t *freeBytes, int64_t *totalBytes);

/* Calibration probe for HBM byte counters (no counterpart in the
 * reference): copies n doubles device-to-device reps times with a plain
 * streaming kernel (umt_copy_probe) and returns the mean ms per copy. */
int umt_hbm_copy_probe(int64_t n, int reps, double *ms);

/* dims[0..6] = nzones, ncorner, nbelem, groups, nangles, levels, items. */
int umt_ctx_dims(const umt_ctx *ctx, int64_t *dims);

/* ------------------------------------------------------------------------
 * Batches of uneven domains. No counterpart in the reference, which sweeps
 * one spatial domain per MPI rank; here many independent problems (each its
 * own mesh, quadrature, group count and cross sections) share the GPU.
 *
 * Level h of a batched sweep holds level h of every domain (plane h of each
 * of its angles in the current angle block), so one kernel launch covers all
 * domains at that depth. A whole source iteration (source update, level
 * launches and phi per angle block, residual per domain) replays as one
 * hipGraph. Groups are stored padded to a multiple of 4 (padding groups
 * carry no flux). Each domain's sweep gives the bits a single-domain context
 * gives for the same groups per lane.
 * ---------------------------------------------------------------------- */
typedef struct umt_batch umt_batch;

/* One domain per problem (umt_problem_tiled), groups[d] groups each,
 * precision 64 or 32 (FP64 phi and residual either way). Starts with 4
 * groups per lane and blocks of up to 32 angles. */
int umt_batch_create(int nDomains, const umt_problem *const *problems, const int *groups,
                     int precision, umt_batch **out);
void umt_batch_destroy(umt_batch *batch);
/* dims[0..5] = domains, unknowns per batched sweep (corners x angles x groups
 * summed), sweep-kernel launches per sweep (levels, or angle blocks when
 * persistent), groups per lane, angles per block, angle blocks. */
int umt_batch_dims(const umt_batch *batch, int64_t *dims);
/* sigt, sigs, q (groups[d], nzones of domain d). */
int umt_batch_set_cross_sections(umt_batch *batch, int domain, const double *sigt,
                                 const double *sigs, const double *q);
/* phi = 0 in every domain. */
int umt_batch_reset(umt_batch *batch);
/* One source-iteration step of every domain without the residual: S from
 * phi, then phi from one sweep. */
int umt_batch_sweep(umt_batch *batch);
int umt_batch_get_phi(const umt_batch *batch, int domain, double *phi); /* (groups[d], C_d) */
/* Source iteration of all domains together until every domain's residual
 * (rt/ConvergenceTest.F90:150-168 semantics) is below tol: iters[d] is the
 * iteration at which domain d first met it, residuals[d] its last residual,
 * total the iterations run. Returns 0 when all converged, 1 otherwise. */
int umt_batch_source_iteration(umt_batch *batch, double tol, int maxIters, int *iters,
                               double *residuals, int *total);
int umt_batch_set_graphs(umt_batch *batch, int on);
/* Groups per lane K (1, 2, 4) and angles per block for the batched sweep. */
int umt_batch_set_layout(umt_batch *batch, int groupsPerLane, int anglesPerBlock);
/* on = 1: every level of an angle block in one persistent launch (a grid
 * the device holds at once walks the levels with a grid barrier between
 * them; same results as one launch per level). A workgroup waits at most
 * maxPolls polls at a barrier (0: about a second; < 0: none, so that a
 * sweep of two or more workgroups fails, to exercise the error path) and
 * then gives up; the sweep then fails with an error and its results are
 * invalid. No
 * counterpart in the reference (its levels are a loop inside one kernel,
 * gpu/GPU_SweepUCBxyz.cu:278-515, for one angle at a time). */
int umt_batch_set_persistent(umt_batch *batch, int on, int maxPolls);
/* Whether the persistent sweep is on, and its workgroups. */
int umt_batch_get_persistent(const umt_batch *batch, int *on, int *blocks);
/* Mean ms of one source iteration (with residuals) over reps, after warmup. */
int umt_batch_time_iterations(umt_batch *batch, int warmup, int reps, double *msPerIteration);
/* Autotuner: times every K in {1, 2, 4} with angles per block in {8, 16, 32,
 * 64, all} (skipping blocks whose psi would not fit in free HBM) with one
 * launch per level, then the persistent sweep at the fastest of those;
 * keeps the fastest and resets phi. */
int umt_batch_autotune(umt_batch *batch, int reps, int *bestK, int *bestAnglesPerBlock,
                       double *msPerIteration);

/* ------------------------------------------------------------------------
 * Learned correction of source iteration. No counterpart in the reference:
 * it accelerates the iteration rt/rtmainsn.F90:167-235 drives, next to GTA.
 *
 * A per-(zone, group) network (10 features in 16 slots -> 64 -> 64 -> 1,
 * tanh; BF16 MFMA with FP32 accumulation, FP32 Adam masters) predicts r, and
 * every `every` sweeps the state becomes phi_half + r (phi_half - phi_prev)
 * per corner. Plain source iteration stays the arbiter: the residual is always
 * that of a plain sweep, a fixed point of plain iteration is left unchanged
 * (the correction scales the sweep's change), and with guard != 0 the
 * corrections stop for the rest of the solve as soon as a corrected state's
 * sweep changes phi more than the sweep the correction followed.
 *
 * Training data come from the method's own iterates: with collection on,
 * source iteration records each state it could correct (features, zone
 * averages, change); umt_ctx_learn_commit turns them into samples whose
 * target is the correction to the context's current (converged) zone
 * averages, (phibar* - phibar_half) / phibar_half, both that and the change
 * scaled by the state's rms change.
 * ---------------------------------------------------------------------- */
typedef struct umt_learner umt_learner;

/* Parameter count: W1 (16 x 64, feature-major), b1 (64), W2 (64 x 64, in-major),
 * b2 (64), w3 (64), b3. */
int umt_learner_num_params(void);
/* New network with Uniform(+-1/sqrt(fan_in)) weights drawn from seed. */
int umt_learner_create(uint64_t seed, umt_learner **out);
void umt_learner_destroy(umt_learner *net);
int umt_learner_get_params(const umt_learner *net, float *params);
/* Sets the FP32 parameters and resets the Adam moments. */
int umt_learner_set_params(umt_learner *net, const float *params);
/* r for n samples of 16 FP32 feature slots each. */
int umt_learner_forward(umt_learner *net, const float *X, int64_t n, float *r);
/* Gradient of (1/n) sum (r a - b)^2 over n samples, and sum (r a - b)^2. */
int umt_learner_gradient(umt_learner *net, const float *X, const float *a, const float *b,
                         int64_t n, float *grad, double *sse);
/* One Adam step (beta 0.9, 0.999, eps 1e-8) on the given batch. */
int umt_learner_step(umt_learner *net, const float *X, const float *a, const float *b, int64_t n,
                     double lr);
int umt_learner_add_samples(umt_learner *net, const float *X, const float *a, const float *b,
                            int64_t n);
int umt_learner_num_samples(const umt_learner *net, int64_t *n);
int umt_learner_get_samples(const umt_learner *net, float *X, float *a, float *b);
int umt_learner_clear_samples(umt_learner *net);
/* Adam over the held samples: epochs passes in seeded random order,
 * mini-batches of `batch`; mean_loss = the last epoch's mean (r a - b)^2. */
int umt_learner_train(umt_learner *net, int epochs, int batch, double lr, uint64_t seed,
                      double *mean_loss);

/* Zone face neighbours, zoneOpp(6, nzones) 1-based with <= 0 on the boundary
 * (mods/Geometry_mod.F90 zoneOpp); umt_ctx_create_from_problem sets them. */
int umt_ctx_set_zone_neighbors(umt_ctx *ctx, const int *zoneOpp);
/* Network used by acceleration mode 2 (NULL detaches), corrections every
 * `every` sweeps, guard as described above. */
int umt_ctx_set_learner(umt_ctx *ctx, umt_learner *net, int every, int guard);
/* on != 0: source iteration records every stride-th state it could correct
 * (in modes 1 and 3 the GTA-corrected states). */
int umt_ctx_learn_collect(umt_ctx *ctx, int on, int stride);
/* Appends the recorded states whose residual is >= min_residual to net's
 * samples, targets against the current zone averages; drops the records. */
int umt_ctx_learn_commit(umt_ctx *ctx, umt_learner *net, double min_residual, int *states);
int umt_ctx_learn_discard(umt_ctx *ctx);
/* The 16 feature slots per (zone, group) for zone averages ph (after a
 * sweep), pp (before it) and dp (the previous sweep's change; NULL: none),
 * all (groups, nzones). X: (16, groups, nzones) FP32. */
int umt_ctx_learn_features(umt_ctx *ctx, const double *ph, const double *pp, const double *dp,
                           float *X);
/* Corrections applied on this context so far, and the iteration at which
 * the last solve stopped correcting (0: it did not). */
int umt_ctx_learn_stats(const umt_ctx *ctx, int64_t *corrections, int *fell_back_at);

#ifdef __cplusplus
}
#endif

#endif /* UMT_SWEEP_H */
