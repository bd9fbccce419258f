/*
 * libumt_amd.so: the 2D cylindrical (RZ) upstream-corner-balance sweep on
 * AMD MI355X (gfx950). C ABI, plain pointers and sizes.
 *
 * The reference sweeps one RZ direction at a time with
 * SweepUCBrz(Set, setID, Groups, Angle, savePsi) (snac/SweepUCBrz.F90:11),
 * called from SetSweep (snac/SetSweep.F90:159-162) for every direction that
 * is not a finishing direction, the directions of each xi level in order
 * because each passes its half-angle flux PsiM to the next. Its GPU form is
 * OpenMP offload (gpu/SweepUCBrz_OMPOL.F90); there is no C entry point for
 * it. This header is the device-resident equivalent: a problem (quad mesh,
 * corner geometry, RZ product quadrature with its angular coefficients, the
 * 2D schedule) and a context that sweeps every direction of every xi level
 * in one pass of level launches and runs steady source iteration.
 *
 * Layouts follow Teton for ndim = 2 (mods/Geometry_mod.F90:23-48): column-
 * major, groups fastest, 1-based index contents, maxcf = 2, four corners per
 * zone. omega(2, A) = (mu, xi), the direction cosines along r and z.
 * Functions return 0 on success; otherwise a nonzero code, with
 * umt_last_error() (include/umt_sweep.h) giving the reason.
 */
#ifndef UMT_RZ_H
#define UMT_RZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct umt_rz_problem umt_rz_problem;

/* nr x nz quad zones on [r0, r1] x [z0, z1] (0 <= r0 < r1), zones r-fastest,
 * interior nodes moved by up to jitter (< 0.5) times the cell size with
 * counter-based draws keyed by seed; RZ product quadrature of npolar x
 * nazimuthal (rt/quadrz.F90:184-268, each 1..4): npolar * 2 xi levels of
 * 2 nazimuthal + 2 directions. */
int umt_rz_problem_structured(int nr, int nz, double r0, double r1, double z0, double z1,
                              double jitter, uint64_t seed, int npolar, int nazimuthal,
                              umt_rz_problem **out);
/* Any quad mesh: points rz(2, npoints) as (r >= 0, z), four node ids per
 * zone quads(4, nzones), 0-based, either orientation (counterclockwise zones
 * are reversed to clockwise, the order of the corner-face conventions). The
 * corner topology is aux/setTetonZone.F90:122-142 with face k = (node k,
 * node k+1) and one vacuum boundary; geometry rt/geometryUCBrz.F90. */
int umt_rz_problem_from_quad_mesh(int npoints, const double *rz, int nzones, const int *quads,
                                  int npolar, int nazimuthal, umt_rz_problem **out);
void umt_rz_problem_destroy(umt_rz_problem *p);

/* dims: nzones, ncorner, nbelem, nangles, npoints, total hyperplanes over
 * directions, xi levels, device sweep levels. */
int umt_rz_problem_dims(const umt_rz_problem *p, int64_t *dims);
/* numCorner(Z), cOffSet(Z), cFP(2,C), cEZ(2,C), CToFace(2,C), zoneOpp(4,Z),
 * faceOpp(4,Z), px(2,C); any output may be NULL. */
int umt_rz_problem_mesh(const umt_rz_problem *p, int *numCorner, int *cOffSet, int *cFP, int *cEZ,
                        int *CToFace, int *zoneOpp, int *faceOpp, double *px);
/* The mesh's points rz(2, npoints) and clockwise quads(4, nzones). */
int umt_rz_problem_points(const umt_rz_problem *p, double *rz, int *quads);
/* A_fp(2,2,C), A_ez(2,2,C), RadiusFP(2,C), RadiusEZ(2,C), Area(C),
 * Volume(C) (r-weighted), VolumeZone(Z). */
int umt_rz_problem_geometry(const umt_rz_problem *p, double *A_fp, double *A_ez, double *RadiusFP,
                            double *RadiusEZ, double *Area, double *Volume, double *VolumeZone);
/* omega(2,A), weight(A) summing to 2 pi, starting(A), finishing(A) flags,
 * angDerivFac(A), quadTauW1(A), quadTauW2(A) (mods/AngleSet_mod.F90:312-347). */
int umt_rz_problem_quadrature(const umt_rz_problem *p, double *omega, double *weight, int *starting,
                              int *finishing, double *angDerivFac, double *quadTauW1,
                              double *quadTauW2);
/* nHyperPlanes(A) (0 for finishing directions, which are not swept),
 * zonesInPlane (concatenated), nextZ(Z,A), nextC(C,A). */
int umt_rz_problem_schedule(const umt_rz_problem *p, int *nHyperPlanes, int *zonesInPlane,
                            int *nextZ, int *nextC);

typedef struct umt_rz_ctx umt_rz_ctx;

/* Device-resident RZ context: geometry, angle table, sweep levels, psi of
 * every direction (G, C + nbelem, A), the half-angle flux of every xi level
 * (G, C, levels), phi (G, C) FP64. precision 64 or 32 (FP32 sweep; phi and
 * the residual stay FP64). */
int umt_rz_ctx_create(const umt_rz_problem *p, int groups, int precision, umt_rz_ctx **out);
void umt_rz_ctx_destroy(umt_rz_ctx *ctx);
/* dims[0..7]: nzones, ncorner, nbelem, groups, nangles, level launches per
 * sweep, (zone, direction) items per sweep, 1 when sweeping in tile levels. */
int umt_rz_ctx_dims(const umt_rz_ctx *ctx, int64_t *dims);
int umt_rz_ctx_set_sigt(umt_rz_ctx *ctx, const double *sigt);      /* (G, Z) */
int umt_rz_ctx_set_source(umt_rz_ctx *ctx, const double *STotal);  /* (G, C) per steradian */
int umt_rz_ctx_set_scattering(umt_rz_ctx *ctx, const double *sigs, const double *q);  /* (G, Z) */
/* One sweep of every direction (SetSweep semantics, snac/SetSweep.F90:90-170):
 * PsiM starts at zero for every xi level, boundaries are vacuum, phi =
 * sum_a w_a psi_a. */
int umt_rz_ctx_sweep(umt_rz_ctx *ctx);
int umt_rz_ctx_get_phi(const umt_rz_ctx *ctx, double *phi);        /* (G, C) */
int umt_rz_ctx_set_phi(umt_rz_ctx *ctx, const double *phi);        /* (G, C) */
/* psi(G, C + nbelem, A): every direction, finishing directions holding the
 * last half-angle flux of their level, exiting boundary fluxes in the tail. */
int umt_rz_ctx_get_psi(const umt_rz_ctx *ctx, double *psi);
/* Steady source iteration: STotal = (sigs phi + q) / (2 pi), sweep, residual
 * as umt_ctx_source_iteration over four-corner zone averages. Returns 0 when
 * converged, 1 when maxIters was reached. */
int umt_rz_ctx_source_iteration(umt_rz_ctx *ctx, double tol, int maxIters, int *iters,
                                double *residual, double *residual_history);
/* Time-dependent sweeps. tau = 1/(c dt) (control/initializeSets.F90:81):
 * the sweep sees Sigt + tau (control/setTotalOpacity.F90) and the source
 * STotal + tau Psi(g, c, Angle) (snac/SweepUCBrz.F90:136), Psi being the
 * previous cycle's angular flux held on the device for every direction.
 * tau = 0 (default) is the steady problem. umt_rz_ctx_save_psi is the
 * cycle's final sweep with savePsi (rt/rtmainsn.F90:238-246,
 * SweepUCBrz.F90:268-276): source from the current phi, one sweep, Psi = its
 * psi (finishing directions: the last half-angle flux), phi = its phi.
 * Psi arrays on the host are Teton's Psi(G, C, A). */
int umt_rz_ctx_set_tau(umt_rz_ctx *ctx, double tau);
int umt_rz_ctx_set_psi_prev(umt_rz_ctx *ctx, const double *Psi);
int umt_rz_ctx_get_psi_prev(const umt_rz_ctx *ctx, double *Psi);
int umt_rz_ctx_save_psi(umt_rz_ctx *ctx);
/* Groups carried per lane (1, 2 or 4, dividing G); default the largest that divides G. */
int umt_rz_ctx_set_groups_per_lane(umt_rz_ctx *ctx, int K);
/* Autotuner: times reps sweeps of the current source with each supported
 * groups-per-lane value and keeps the fastest. msPerK[0..2] receives the
 * mean ms per sweep for K = 1, 2, 4 (-1 when unsupported); may be NULL.
 * Overwrites psi and phi. No counterpart in the reference. */
int umt_rz_ctx_autotune(umt_rz_ctx *ctx, int reps, int *bestK, double *msPerK);
/* on != 0: sweep in tile levels, one wave walking the zone levels of a tile
 * of about 4 x 4 zones for one direction between barriers, so a launch
 * covers several zone levels; 0 (default): one launch per zone level, which
 * is faster on MI355X today. Both give the same bits. Fails when the problem
 * has no tile schedule (some direction's tile graph has a cycle).
 * No counterpart in the reference. */
int umt_rz_ctx_set_tiles(umt_rz_ctx *ctx, int on);
/* Replay each sweep from a captured hipGraph (default on; kernel nodes
 * only, rebuilt per groups-per-lane choice) or launch it eagerly. */
int umt_rz_ctx_set_graphs(umt_rz_ctx *ctx, int on);
/* Mean time of one sweep (ms) and of its level kernels alone (ms) over reps
 * sweeps after warmup, HIP events on the context's stream (with graphs the
 * level launches are replayed back to back from a graph of their own;
 * without, each level is bracketed by events and the durations summed). */
int umt_rz_ctx_time_sweeps(umt_rz_ctx *ctx, int warmup, int reps, double *msSweep,
                           double *msLevelKernels);

#ifdef __cplusplus
}
#endif

#endif /* UMT_RZ_H */
