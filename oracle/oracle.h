/*
 * ORACLE — test infrastructure only. CPU restatement of the reference
 * (LLNL UMT / Teton 5.3.0) UCB xyz sweep path in plain C, FP64.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the CPU baseline; the product library
 * never links or calls it.
 *
 * Layout matches Teton: column-major shapes (first index fastest, groups
 * fastest), 1-based index contents. Each function cites the reference
 * routine it restates (paths relative to src/teton/).
 *
 * How each function is pinned (the reference has no tests or golden vectors
 * for this path, and its Fortran cannot be built here: SweepUCBxyz.F90 uses
 * Size_mod, which pulls in Conduit, MPI and a CMake-generated module):
 *   oracle_geometry          closure, exact volumes, equal-and-opposite areas,
 *                            box-zone closed forms (tests/test_oracle_geometry.py)
 *   oracle_quadrature        QRS45 table fixture parsed from the reference
 *                            text, sum w = 4 pi, octant symmetry
 *   oracle_schedule_angle    brute-force longest-path layering, brute-force
 *                            lexicographically-smallest corner order
 *   oracle_schedule_angle_cyc  on cyclic meshes: every psi read before it is
 *                            computed is on the cycle list, cyclic zones have
 *                            all corners listed, every zone once; the lagged
 *                            sweep's fixed point solves the coupled corner
 *                            equations of all zones (dense solve)
 *   oracle_sweep_angle       zone balance, constant solution, thick limit,
 *                            corner balance with the grey-form coupling,
 *                            analytic pure-absorber convergence
 *   oracle_source_iteration  infinite-medium fixed point, global balance
 */
#ifndef UMT_ORACLE_H
#define UMT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* rt/geometryUCBxyz.F90:63-181 with mods/Geometry_mod.F90:344-441 centres. */
int oracle_geometry(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                    const int *zoneFaces, const int *nCFaces, const int *cFP, const int *cEZ,
                    const int *CToFace, const int *zoneOpp, const double *px, double *A_fp,
                    double *A_ez, double *Volume);

/* rt/quadProduct.F90:52-182 + rt/rtquad.F90:95-105 (polar axis 1..3) from
 * the tabulated polar (cosTheta, weightTheta) and azimuthal (cosPhi,
 * weightPhi) values of one order each. Writes omega(3, 8*np*na), weight. */
int oracle_quadrature(int npolar, int nazimuthal, int polaraxis, const double *cosTheta,
                      const double *weightTheta, const double *cosPhi, const double *weightPhi,
                      double *omega, double *weight);

/* snac/snneed.F90:95-191, snac/findseeds.F90:58-66,
 * snac/getDownStreamData.F90:88-149, snac/snnext.F90:103-195 for one angle.
 * Outputs nextZ(Z), nextC(C), zonesInPlane(<=Z), *nHyperPlanes.
 * Returns 0, or 2 when the angle's schedule has a cycle list (use
 * oracle_schedule_angle_cyc for those), or another nonzero code where the
 * reference stops with a fatal error. */
int oracle_schedule_angle(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                          const int *zoneFaces, const int *nCFaces, const int *cEZ,
                          const int *CToFace, const int *zoneOpp, const int *faceOpp,
                          const double *A_fp, const double *A_ez, const double *omega,
                          int *nextZ, int *nextC, int *zonesInPlane, int *nHyperPlanes);

/* The same with corner faces whose |Omega.A| <= parTol * |A|_1 taken as
 * parallel to Omega (no dependency): directions lying in mesh faces, as the
 * S2 grey-acceleration directions do on the tile's diagonal faces. parTol = 0
 * is oracle_schedule_angle. */
int oracle_schedule_angle_tol(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                              const int *zoneFaces, const int *nCFaces, const int *cEZ,
                              const int *CToFace, const int *zoneOpp, const int *faceOpp,
                              const double *A_fp, const double *A_ez, const double *omega,
                              double parTol, int *nextZ, int *nextC, int *zonesInPlane,
                              int *nHyperPlanes);

/* The full schedule of one angle with the reference's cycle handling:
 * snneed's mixed-sign faces, findseeds' forced seed (snac/findseeds.F90:68-114),
 * getDownStreamData's cyclic zones (nextZ < 0, natural corner order,
 * snac/fixZone.F90) and cyclebreaker's Tarjan search (snac/cyclebreaker.F90,
 * snac/sccsearch.F90), all appending the corners read before they are
 * computed to cycleList (1-based, capacity ncorner; *numCycles entries).
 * Returns 0, or nonzero where the reference stops with a fatal error. */
int oracle_schedule_angle_cyc(int nzones, int ncorner, const int *numCorner, const int *cOffSet,
                              const int *zoneFaces, const int *nCFaces, const int *cEZ,
                              const int *CToFace, const int *zoneOpp, const int *faceOpp,
                              const double *A_fp, const double *A_ez, const double *omega,
                              double parTol, int *nextZ, int *nextC, int *zonesInPlane,
                              int *nHyperPlanes, int *cycleList, int *numCycles);

/* snac/SweepUCBxyz.F90:82-318 for one angle, in nextZ / nextC order.
 * Psi1(G, C+nbelem) work array; PsiB(G, nbelem) of this angle (in: incident,
 * out: exiting); Phi(G, C) accumulated with quadwt; tau*Psi term used when
 * tau != 0 (Psi(G,C) of this angle, may be NULL when tau == 0). */
void oracle_sweep_angle(int nzones, int ncorner, int nbelem, int G, const int *numCorner,
                        const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                        const double *A_fp, const double *A_ez, const double *Volume,
                        const double *STotal, const double *Sigt, double tau, const double *Psi,
                        const double *omega, double quadwt, int nHyperPlanes,
                        const int *zonesInPlane, const int *nextZ, const int *nextC,
                        double *Psi1, double *PsiB, double *Phi);

/* The same with this angle's cycle list (numCycles 1-based corners) and
 * their lagged psi cyclePsi(G, numCycles), loaded into Psi1 before the
 * sweep and stored after it (control/constructDynMemory.F90:118-189). */
void oracle_sweep_angle_cyc(int nzones, int ncorner, int nbelem, int G, const int *numCorner,
                            const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                            const double *A_fp, const double *A_ez, const double *Volume,
                            const double *STotal, const double *Sigt, double tau,
                            const double *Psi, const double *omega, double quadwt,
                            int nHyperPlanes, const int *zonesInPlane, const int *nextZ,
                            const int *nextC, int numCycles, const int *cycleList,
                            double *cyclePsi, double *Psi1, double *PsiB, double *Phi);

/* snac/SetSweep.F90:100-170 for a single domain with vacuum boundaries:
 * Phi = 0, then every angle swept (OpenMP over angles when nthreads > 1,
 * per-thread Phi partials summed in angle-block order). PsiB(G, nbelem, A)
 * holds incident fluxes on entry, exiting ones on return. Schedule arrays
 * are per angle: nHyperPlanes(A), zonesInPlane concatenated, nextZ(Z,A),
 * nextC(C,A). */
void oracle_sweep_all(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                      const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                      const double *A_fp, const double *A_ez, const double *Volume,
                      const double *STotal, const double *Sigt, const double *omega,
                      const double *weight, const int *nHyperPlanes, const int *zonesInPlane,
                      const int *nextZ, const int *nextC, int nthreads, double *PsiB,
                      double *Phi);

/* The same with corner faces whose |Omega.A| <= parTol * |A|_1 taken as
 * parallel to Omega in the sweep as in the schedule (oracle_schedule_angle_tol):
 * their Omega.A is exactly 0, so no upwind value is read across them and
 * they do not select the opposite-face branch of the EZ coupling. */
void oracle_sweep_all_tol(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                          const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                          const double *A_fp, const double *A_ez, const double *Volume,
                          const double *STotal, const double *Sigt, const double *omega,
                          const double *weight, const int *nHyperPlanes, const int *zonesInPlane,
                          const int *nextZ, const int *nextC, int nthreads, double parTol,
                          double *PsiB, double *Phi);

/* The same with cycle lists: numCycles(A) and cycleList (angle by angle, as
 * rt/rtorder.F90:62-90 concatenates them); cyclePsi(G, total) holds each
 * cycle corner's psi from the previous sweep of its angle, loaded into Psi1
 * before the angle's sweep and stored after it
 * (control/constructDynMemory.F90:118-189). numCycles == NULL: no cycles. */
void oracle_sweep_all_cyc(int nzones, int ncorner, int nbelem, int G, int A, const int *numCorner,
                          const int *cOffSet, const int *nCFaces, const int *cFP, const int *cEZ,
                          const double *A_fp, const double *A_ez, const double *Volume,
                          const double *STotal, const double *Sigt, const double *omega,
                          const double *weight, const int *nHyperPlanes, const int *zonesInPlane,
                          const int *nextZ, const int *nextC, int nthreads, double parTol,
                          const int *numCycles, const int *cycleList, double *cyclePsi,
                          double *PsiB, double *Phi);

/* Steady source iteration (rt/rtmainsn.F90:167-235, rt/LinearSolver.F90:55-81,
 * rt/ConvergenceTest.F90:150-168 semantics): STotal = (sigs*phi + q)/(4 pi),
 * sweep all angles, residual = max over (zone,group) with
 * phibar > 1e-6*max(phibar) of |phibar_new - phibar_old|/phibar_new.
 * Phi(G,C) holds the starting guess and the result. Runs until
 * residual < tol or maxIters. Returns iterations done; *residual is the last. */
int oracle_source_iteration(int nzones, int ncorner, int nbelem, int G, int A,
                            const int *numCorner, const int *cOffSet, const int *nCFaces,
                            const int *cFP, const int *cEZ, const double *A_fp,
                            const double *A_ez, const double *Volume, const double *VolumeZone,
                            const double *Sigt, const double *sigs, const double *q,
                            const double *omega, const double *weight, const int *nHyperPlanes,
                            const int *zonesInPlane, const int *nextZ, const int *nextC,
                            int nthreads, double tol, int maxIters, double *Phi,
                            double *residual, double *residual_history);

/* The same with cycle lists (as oracle_sweep_all_cyc); cyclePsi is carried
 * from sweep to sweep (its entry value is the first sweep's). */
int oracle_source_iteration_cyc(int nzones, int ncorner, int nbelem, int G, int A,
                                const int *numCorner, const int *cOffSet, const int *nCFaces,
                                const int *cFP, const int *cEZ, const double *A_fp,
                                const double *A_ez, const double *Volume, const double *VolumeZone,
                                const double *Sigt, const double *sigs, const double *q,
                                const double *omega, const double *weight, const int *nHyperPlanes,
                                const int *zonesInPlane, const int *nextZ, const int *nextC,
                                const int *numCycles, const int *cycleList, double *cyclePsi,
                                int nthreads, double tol, int maxIters, double *Phi,
                                double *residual, double *residual_history);

#ifdef __cplusplus
}
#endif

#endif /* UMT_ORACLE_H */
