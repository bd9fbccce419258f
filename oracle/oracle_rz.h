/*
 * ORACLE — test infrastructure only. CPU restatement of the reference
 * (LLNL UMT / Teton 5.3.0) 2D cylindrical (RZ) upstream-corner-balance sweep
 * path in plain C, FP64. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, as the checker or the CPU baseline; the
 * product library never links or calls it.
 *
 * Layout as Teton's (ndim = 2, maxcf = 2, quad zones): column-major shapes,
 * groups fastest, 1-based index contents. px(2, C) = (r, z) of each corner's
 * point; A_fp / A_ez (2, 2, C); RadiusFP / RadiusEZ (2, C); omega(2, A) =
 * (mu, xi).
 *
 * How each function is pinned (tests/test_oracle_rz.py; the reference has no
 * tests or fixtures for this path and cannot be built here):
 *   oracle_rz_geometry        closure of every corner, corner areas summing to
 *                             the shoelace zone area, volumes summing to the
 *                             exact r-weighted zone volume (Pappus), and the
 *                             divergence identity sum_f R_f Omega.A_f = mu Area
 *   oracle_rz_quadrature      the QRS45 fixture tables, sum w = 2 pi, mu rising
 *                             in each xi level, starting/finishing flags
 *   oracle_rz_angle_coef      closed forms: alpha recursion ends at zero, tau
 *                             within (0, 1], the cell edges of each level
 *   oracle_rz_schedule_angle  brute-force longest-path layering, brute-force
 *                             smallest corner order
 *   oracle_rz_sweep_*         the constant isotropic solution is exact through
 *                             every angle of a level (angular redistribution
 *                             included); global particle balance
 *   oracle_rz_source_iteration  infinite-medium fixed point, global balance
 */
#ifndef UMT_ORACLE_RZ_H
#define UMT_ORACLE_RZ_H

#ifdef __cplusplus
extern "C" {
#endif

/* rt/geometryUCBrz.F90:45-157 with mods/Geometry_mod.F90:344-378 zone centres. */
int oracle_rz_geometry(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                       const int *cFP, const int *cEZ, const double *px, double *A_fp,
                       double *A_ez, double *RadiusFP, double *RadiusEZ, double *Area,
                       double *Volume);

/* rt/quadrz.F90:184-268 (product quadrature) and rt/rtquad.F90:80-123 from
 * one order's polar (cosTheta, weightTheta) and RZ azimuthal (cosPhiRZ,
 * weightPhiRZ) tables. A = npolar * 2 * (2 nazimuthal + 2) directions:
 * omega(2, A), weight(A) summing to 2 pi, starting(A), finishing(A). */
int oracle_rz_quadrature(int npolar, int nazimuthal, const double *cosTheta,
                         const double *weightTheta, const double *cosPhiRZ,
                         const double *weightPhiRZ, double *omega, double *weight, int *starting,
                         int *finishing);

/* rt/AngleCoef2D.F90:30-100 and mods/AngleSet_mod.F90:312-347: the angular
 * derivative factor and the half-angle weights of every direction. Returns
 * nonzero where the reference stops with a fatal error. */
int oracle_rz_angle_coef(int A, const double *omega, const double *weight, const int *starting,
                         const int *finishing, double *angDerivFac, double *quadTauW1,
                         double *quadTauW2);

/* snac/snneed.F90:62-93 (2D), findseeds.F90:58-66, getDownStreamData.F90:88-149
 * (two corner faces), snnext.F90:103-195 for one direction omega(2).
 * Returns 0, or nonzero for a schedule with cycles (not restated in 2D). */
int oracle_rz_schedule_angle(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                             const int *zoneFaces, const int *cEZ, const int *CToFace,
                             const int *zoneOpp, const int *faceOpp, const double *A_fp,
                             const double *A_ez, const double *omega, int *nextZ, int *nextC,
                             int *zonesInPlane, int *nHyperPlanes);

/* snac/SweepUCBrz.F90:105-276 for one direction: Psi1(G, C+nbelem) work
 * array, PsiM(G, C) half-angle flux carried from the previous direction of
 * the xi level (in) to the next (out), PsiB(G, nbelem) of this direction,
 * PsiBnext (this direction's exiting elements for a finishing next
 * direction, or NULL), Phi(G, C) accumulated, PsiNext (G, C) receives PsiM
 * for a finishing next direction (or NULL). */
void oracle_rz_sweep_angle(int nzones, int ncorner, int nbelem, int G, const int *numCorner,
                           const int *cOffSet, const int *cFP, const int *cEZ, const double *A_fp,
                           const double *A_ez, const double *RadiusFP, const double *RadiusEZ,
                           const double *Area, const double *Volume, const double *STotal,
                           const double *Sigt, double tau, const double *Psi,
                           const double *omega, double quadwt, double fac, double quadTauW1,
                           double quadTauW2, int startingDirection, int nHyperPlanes,
                           const int *zonesInPlane, const int *nextZ, const int *nextC,
                           double *Psi1, double *PsiM, double *PsiB, double *PsiBnext,
                           double *Phi, double *PsiNext);

/* snac/SetSweep.F90:90-170 for one RZ domain with vacuum boundaries (PsiB
 * holds incident fluxes on entry, exiting ones on return; PsiM and Phi start
 * at zero). Schedule arrays per direction; finishing directions are not
 * swept. Psi (G, C, A) may be NULL. */
void oracle_rz_sweep_all(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                         const int *cOffSet, const int *cFP, const int *cEZ, const double *A_fp,
                         const double *A_ez, const double *RadiusFP, const double *RadiusEZ,
                         const double *Area, const double *Volume, const double *STotal,
                         const double *Sigt, const double *omega, const double *weight,
                         const int *starting, const int *finishing, const double *angDerivFac,
                         const double *quadTauW1, const double *quadTauW2,
                         const int *nHyperPlanes, const int *zonesInPlane, const int *nextZ,
                         const int *nextC, double tau, const double *PsiPrev, int nthreads,
                         double *PsiB, double *Phi, double *Psi);

/* Steady source iteration in RZ (STotal = (sigs phi + q) / (2 pi)), residual
 * as oracle_source_iteration over four-corner zone averages. */
int oracle_rz_source_iteration(int nzones, int ncorner, int nbelem, int G, int A,
                               const int *numCorner, const int *cOffSet, const int *cFP,
                               const int *cEZ, const double *A_fp, const double *A_ez,
                               const double *RadiusFP, const double *RadiusEZ, const double *Area,
                               const double *Volume, const double *VolumeZone, const double *Sigt,
                               const double *sigs, const double *q, const double *omega,
                               const double *weight, const int *starting, const int *finishing,
                               const double *angDerivFac, const double *quadTauW1,
                               const double *quadTauW2, const int *nHyperPlanes,
                               const int *zonesInPlane, const int *nextZ, const int *nextC,
                               int nthreads, double tol, int maxIters, double *Phi,
                               double *residual, double *residual_history);

#ifdef __cplusplus
}
#endif

#endif /* UMT_ORACLE_RZ_H */
