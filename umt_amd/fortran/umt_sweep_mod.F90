!***********************************************************************
!  umt_sweep_mod - Fortran host binding of the MI355X UCB xyz sweep
!                  library (libumt_amd.so, C-ABI in include/umt_sweep.h).
!
!  The drop-in entry point gpu_sweepucbxyz keeps the argument list the
!  reference binds in src/teton/gpu/SweepUCBxyzToGPU.F90:43-91 (scalars by
!  reference, arrays by first element), so a Teton build links this library
!  in place of the CUDA object without touching that file. The remaining
!  interfaces expose the device-resident context: a Fortran host hands over
!  its Teton-layout arrays once (Geometry_mod / AngleSet_mod names and
!  shapes) and then sweeps or runs source iteration on the GPU.
!***********************************************************************
module umt_sweep_mod

  use iso_c_binding

  implicit none

  private

  public :: gpu_sweepucbxyz, gpu_streamsynchronize, gpu_devicesynchronize, &
            gpu_sweepucbxyz_status, umt_problem_tiled, umt_problem_destroy, &
            umt_problem_dims, umt_problem_mesh, umt_problem_geometry,       &
            umt_problem_quadrature, umt_problem_schedule,                   &
            umt_draw_cross_sections, umt_ctx_create, umt_ctx_destroy,       &
            umt_ctx_set_sigt, umt_ctx_set_source, umt_ctx_set_scattering,   &
            umt_ctx_sweep, umt_ctx_get_phi, umt_ctx_source_iteration,       &
            umt_ctx_autotune, umt_error_message, umt_ctx_create_from_problem, &
            umt_ctx_set_phi, umt_ctx_copy_phi, umt_ctx_set_acceleration,      &
            umt_ctx_gta_stats, umt_ctx_set_graphs, umt_ctx_set_angles_per_block, &
            umt_ctx_set_stall_window, umt_ctx_set_zone_neighbors,             &
            umt_learner_create, umt_learner_destroy, umt_learner_train,       &
            umt_ctx_set_learner, umt_ctx_learn_collect, umt_ctx_learn_commit, &
            umt_ctx_learn_stats, umt_problem_cycles, umt_ctx_set_cycles,      &
            umt_problem_from_hex_mesh, umt_problem_gta, umt_ctx_enable_gta,   &
            umt_ctx_dims, umt_ctx_get_psi, umt_ctx_set_groups_per_lane,       &
            umt_ctx_get_groups_per_lane, umt_ctx_get_angles_per_block,        &
            umt_ctx_sweep_n, umt_ctx_time_sweeps, umt_ctx_set_tau,            &
            umt_ctx_set_psi_prev, umt_ctx_get_psi_prev, umt_ctx_save_psi,     &
            umt_batch_create, umt_batch_destroy, umt_batch_dims,              &
            umt_batch_set_cross_sections, umt_batch_reset, umt_batch_sweep,   &
            umt_batch_get_phi, umt_batch_source_iteration,                    &
            umt_batch_set_graphs, umt_batch_set_layout,                       &
            umt_batch_time_iterations, umt_batch_autotune,                    &
            umt_ctx_set_gta_solver, umt_ctx_set_grey_controls, umt_ctx_set_gta_options

  interface

!   gpu/GPU_SweepUCBxyz.cu:532-572 argument list
    subroutine gpu_sweepucbxyz(Angle, nHyperPlanes, nZonesInPlane, nextZ, nextC, STotal, tau, &
                               Psi, Groups, Volume, Sigt, nCFacesArray, ndim, maxcf, ncorner,  &
                               A_fp, omega, cFP, Psi1, nbelem, A_ez, cEZ, NumAngles, quadwt,   &
                               Phi, PsiB, maxCorner, mem0solve1, streamId, totalStreams,       &
                               savePsi, numCycles, cycleOffSet, cyclePsi, cycleList, b0,       &
                               nBdyElem, PsiBMref, Mref, Geom_numCorner, Geom_cOffSet)         &
                               bind(c, name="gpu_sweepucbxyz")
      import :: c_int, c_double
      integer(c_int) :: Angle, nHyperPlanes, nZonesInPlane(*), nextZ(*), nextC(*)
      real(c_double) :: STotal(*), tau, Psi(*)
      integer(c_int) :: Groups
      real(c_double) :: Volume(*), Sigt(*)
      integer(c_int) :: nCFacesArray(*), ndim, maxcf, ncorner
      real(c_double) :: A_fp(*), omega(*)
      integer(c_int) :: cFP(*)
      real(c_double) :: Psi1(*)
      integer(c_int) :: nbelem
      real(c_double) :: A_ez(*)
      integer(c_int) :: cEZ(*), NumAngles
      real(c_double) :: quadwt, Phi(*), PsiB(*)
      integer(c_int) :: maxCorner, mem0solve1, streamId, totalStreams, savePsi, numCycles
      integer(c_int) :: cycleOffSet
      real(c_double) :: cyclePsi(*)
      integer(c_int) :: cycleList(*), b0, nBdyElem
      real(c_double) :: PsiBMref(*)
      integer(c_int) :: Mref, Geom_numCorner(*), Geom_cOffSet(*)
    end subroutine gpu_sweepucbxyz

!   gpu/GPU_SweepUCBxyz.cu:520-528
    subroutine gpu_streamsynchronize(streamId) bind(c, name="gpu_streamsynchronize")
      import :: c_int
      integer(c_int) :: streamId
    end subroutine gpu_streamsynchronize

    subroutine gpu_devicesynchronize() bind(c, name="gpu_devicesynchronize")
    end subroutine gpu_devicesynchronize

    integer(c_int) function gpu_sweepucbxyz_status() bind(c, name="gpu_sweepucbxyz_status")
      import :: c_int
    end function gpu_sweepucbxyz_status

    type(c_ptr) function umt_last_error() bind(c, name="umt_last_error")
      import :: c_ptr
    end function umt_last_error

    integer(c_int) function umt_problem_tiled(nx, ny, nz, npolar, nazimuthal, kdtree, prob) &
        bind(c, name="umt_problem_tiled")
      import :: c_int, c_ptr
      integer(c_int), value :: nx, ny, nz, npolar, nazimuthal, kdtree
      type(c_ptr) :: prob
    end function umt_problem_tiled

    subroutine umt_problem_destroy(prob) bind(c, name="umt_problem_destroy")
      import :: c_ptr
      type(c_ptr), value :: prob
    end subroutine umt_problem_destroy

    integer(c_int) function umt_problem_dims(prob, dims) bind(c, name="umt_problem_dims")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: prob
      integer(c_int64_t) :: dims(8)
    end function umt_problem_dims

    integer(c_int) function umt_problem_mesh(prob, numCorner, cOffSet, nCFaces, cFP, cEZ, &
                                             CToFace, zoneOpp, faceOpp, px)              &
        bind(c, name="umt_problem_mesh")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: prob
      integer(c_int) :: numCorner(*), cOffSet(*), nCFaces(*), cFP(*), cEZ(*), CToFace(*)
      integer(c_int) :: zoneOpp(*), faceOpp(*)
      real(c_double) :: px(*)
    end function umt_problem_mesh

    integer(c_int) function umt_problem_geometry(prob, A_fp, A_ez, Volume, VolumeZone) &
        bind(c, name="umt_problem_geometry")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: prob
      real(c_double) :: A_fp(*), A_ez(*), Volume(*), VolumeZone(*)
    end function umt_problem_geometry

    integer(c_int) function umt_problem_quadrature(prob, omega, weight) &
        bind(c, name="umt_problem_quadrature")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: prob
      real(c_double) :: omega(*), weight(*)
    end function umt_problem_quadrature

    integer(c_int) function umt_problem_schedule(prob, nHyperPlanes, zonesInPlane, nextZ, nextC) &
        bind(c, name="umt_problem_schedule")
      import :: c_int, c_ptr
      type(c_ptr), value :: prob
      integer(c_int) :: nHyperPlanes(*), zonesInPlane(*), nextZ(*), nextC(*)
    end function umt_problem_schedule

!   ASet%numCycles(A) and ASet%cycleList (1-based corners, angle by angle);
!   pass totalCycles alone first (the arrays may be c_null_ptr via the
!   c_ptr form below) to size cycleList.
    integer(c_int) function umt_problem_cycles(prob, totalCycles, numCycles, cycleList) &
        bind(c, name="umt_problem_cycles")
      import :: c_int, c_ptr
      type(c_ptr), value :: prob
      integer(c_int) :: totalCycles
      type(c_ptr), value :: numCycles, cycleList
    end function umt_problem_cycles

    integer(c_int) function umt_draw_cross_sections(seed, nzones, groups, sigt_lo, sigt_hi, &
                                                    c_lo, c_hi, q_lo, q_hi, sigt, sigs, q)  &
        bind(c, name="umt_draw_cross_sections")
      import :: c_int, c_double, c_int64_t
      integer(c_int64_t), value :: seed
      integer(c_int), value :: nzones, groups
      real(c_double), value :: sigt_lo, sigt_hi, c_lo, c_hi, q_lo, q_hi
      real(c_double) :: sigt(*), sigs(*), q(*)
    end function umt_draw_cross_sections

    integer(c_int) function umt_ctx_create(nzones, ncorner, nbelem, groups, nangles, precision, &
                                           numCorner, cOffSet, nCFaces, cFP, cEZ, A_fp, A_ez,   &
                                           Volume, VolumeZone, omega, weight, nHyperPlanes,     &
                                           zonesInPlane, nextZ, ctx)                            &
        bind(c, name="umt_ctx_create")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nzones, ncorner, nbelem, groups, nangles, precision
      integer(c_int) :: numCorner(*), cOffSet(*), nCFaces(*), cFP(*), cEZ(*)
      real(c_double) :: A_fp(*), A_ez(*), Volume(*), VolumeZone(*), omega(*), weight(*)
      integer(c_int) :: nHyperPlanes(*), zonesInPlane(*), nextZ(*)
      type(c_ptr) :: ctx
    end function umt_ctx_create

    subroutine umt_ctx_destroy(ctx) bind(c, name="umt_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine umt_ctx_destroy

    integer(c_int) function umt_ctx_set_sigt(ctx, sigt) bind(c, name="umt_ctx_set_sigt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: sigt(*)
    end function umt_ctx_set_sigt

    integer(c_int) function umt_ctx_set_source(ctx, STotal) bind(c, name="umt_ctx_set_source")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: STotal(*)
    end function umt_ctx_set_source

    integer(c_int) function umt_ctx_set_scattering(ctx, sigs, q) &
        bind(c, name="umt_ctx_set_scattering")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: sigs(*), q(*)
    end function umt_ctx_set_scattering

    integer(c_int) function umt_ctx_sweep(ctx) bind(c, name="umt_ctx_sweep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function umt_ctx_sweep

    integer(c_int) function umt_ctx_get_phi(ctx, phi) bind(c, name="umt_ctx_get_phi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: phi(*)
    end function umt_ctx_get_phi

    integer(c_int) function umt_ctx_source_iteration(ctx, tol, maxIters, iters, residual, hist) &
        bind(c, name="umt_ctx_source_iteration")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tol
      integer(c_int), value :: maxIters
      integer(c_int) :: iters
      real(c_double) :: residual, hist(*)
    end function umt_ctx_source_iteration

    integer(c_int) function umt_ctx_autotune(ctx, reps, bestK, msPerK) &
        bind(c, name="umt_ctx_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      integer(c_int) :: bestK
      real(c_double) :: msPerK(4)
    end function umt_ctx_autotune

    integer(c_int) function umt_ctx_create_from_problem(prob, groups, precision, ctx) &
        bind(c, name="umt_ctx_create_from_problem")
      import :: c_int, c_ptr
      type(c_ptr), value :: prob
      integer(c_int), value :: groups, precision
      type(c_ptr) :: ctx
    end function umt_ctx_create_from_problem

    integer(c_int) function umt_ctx_set_phi(ctx, phi) bind(c, name="umt_ctx_set_phi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: phi(*)
    end function umt_ctx_set_phi

    integer(c_int) function umt_ctx_copy_phi(dst, src) bind(c, name="umt_ctx_copy_phi")
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
    end function umt_ctx_copy_phi

!   The schedule's cycle lists (ASet%numCycles, ASet%cycleList) for meshes
!   whose sweeps have dependency cycles.
    integer(c_int) function umt_ctx_set_cycles(ctx, numCycles, cycleList) &
        bind(c, name="umt_ctx_set_cycles")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: numCycles(*), cycleList(*)
    end function umt_ctx_set_cycles

!   mode 0: plain source iteration, 1: GTA, 2: learned correction
    integer(c_int) function umt_ctx_set_acceleration(ctx, mode) &
        bind(c, name="umt_ctx_set_acceleration")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function umt_ctx_set_acceleration

!   GTA kernel 1 (grey + stretched-TSA sweeps) or 2 (transfer matrices, the
!   reference's Size%useNewGTASolver); resets the hard grey-iteration maximum
!   to the kernel's default (off for 1, on for 2)
    integer(c_int) function umt_ctx_set_gta_solver(ctx, kernel) &
        bind(c, name="umt_ctx_set_gta_solver")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: kernel
    end function umt_ctx_set_gta_solver

!   grey_reltol and grey_max_sweeps (teton_adjust_grey_reltol / _maxits)
    integer(c_int) function umt_ctx_set_grey_controls(ctx, relTol, maxIts) &
        bind(c, name="umt_ctx_set_grey_controls")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: relTol
      integer(c_int), value :: maxIts
    end function umt_ctx_set_grey_controls

!   enforceHardGTAIterMax (teton_setgtaoptions)
    integer(c_int) function umt_ctx_set_gta_options(ctx, enforceHardIterMax) &
        bind(c, name="umt_ctx_set_gta_options")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enforceHardIterMax
    end function umt_ctx_set_gta_options

    integer(c_int) function umt_ctx_gta_stats(ctx, greyIterations, greySweeps) &
        bind(c, name="umt_ctx_gta_stats")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: greyIterations, greySweeps
    end function umt_ctx_gta_stats

    integer(c_int) function umt_ctx_set_graphs(ctx, on) bind(c, name="umt_ctx_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function umt_ctx_set_graphs

    integer(c_int) function umt_ctx_set_angles_per_block(ctx, n) &
        bind(c, name="umt_ctx_set_angles_per_block")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
    end function umt_ctx_set_angles_per_block

    integer(c_int) function umt_ctx_set_stall_window(ctx, window) &
        bind(c, name="umt_ctx_set_stall_window")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: window
    end function umt_ctx_set_stall_window

!   zoneOpp(6, nzones) of mods/Geometry_mod.F90 (1-based, <= 0 on the boundary)
    integer(c_int) function umt_ctx_set_zone_neighbors(ctx, zoneOpp) &
        bind(c, name="umt_ctx_set_zone_neighbors")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: zoneOpp(*)
    end function umt_ctx_set_zone_neighbors

    integer(c_int) function umt_learner_create(seed, net) bind(c, name="umt_learner_create")
      import :: c_int, c_ptr, c_int64_t
      integer(c_int64_t), value :: seed
      type(c_ptr) :: net
    end function umt_learner_create

    subroutine umt_learner_destroy(net) bind(c, name="umt_learner_destroy")
      import :: c_ptr
      type(c_ptr), value :: net
    end subroutine umt_learner_destroy

    integer(c_int) function umt_learner_train(net, epochs, batch, lr, seed, meanLoss) &
        bind(c, name="umt_learner_train")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: net
      integer(c_int), value :: epochs, batch
      real(c_double), value :: lr
      integer(c_int64_t), value :: seed
      real(c_double) :: meanLoss
    end function umt_learner_train

    integer(c_int) function umt_ctx_set_learner(ctx, net, every, guard) &
        bind(c, name="umt_ctx_set_learner")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, net
      integer(c_int), value :: every, guard
    end function umt_ctx_set_learner

    integer(c_int) function umt_ctx_learn_collect(ctx, on, stride) &
        bind(c, name="umt_ctx_learn_collect")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on, stride
    end function umt_ctx_learn_collect

    integer(c_int) function umt_ctx_learn_commit(ctx, net, minResidual, states) &
        bind(c, name="umt_ctx_learn_commit")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, net
      real(c_double), value :: minResidual
      integer(c_int) :: states
    end function umt_ctx_learn_commit

    integer(c_int) function umt_ctx_learn_stats(ctx, corrections, fellBackAt) &
        bind(c, name="umt_ctx_learn_stats")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: corrections
      integer(c_int) :: fellBackAt
    end function umt_ctx_learn_stats

!   Any hex mesh: coords(3, npoints), connectivity(8, nzones) 0-based in the
!   Blueprint / VTK hexahedron node order (mesh ingestion, mesh.cpp).
    integer(c_int) function umt_problem_from_hex_mesh(npoints, coords, nzones, connectivity, &
                                                      npolar, nazimuthal, kdtree, prob)      &
        bind(c, name="umt_problem_from_hex_mesh")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: npoints, nzones, npolar, nazimuthal, kdtree
      real(c_double) :: coords(*)
      integer(c_int) :: connectivity(*)
      type(c_ptr) :: prob
    end function umt_problem_from_hex_mesh

!   The S2 grey-acceleration directions and their schedule.
    integer(c_int) function umt_problem_gta(prob, nangles, omega, weight, nHyperPlanes, &
                                            zonesInPlane, nextZ) bind(c, name="umt_problem_gta")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: prob
      integer(c_int) :: nangles
      real(c_double) :: omega(*), weight(*)
      integer(c_int) :: nHyperPlanes(*), zonesInPlane(*), nextZ(*)
    end function umt_problem_gta

    integer(c_int) function umt_ctx_enable_gta(ctx, nangles, omega, weight, nHyperPlanes, &
                                               zonesInPlane, nextZ) bind(c, name="umt_ctx_enable_gta")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nangles
      real(c_double) :: omega(*), weight(*)
      integer(c_int) :: nHyperPlanes(*), zonesInPlane(*), nextZ(*)
    end function umt_ctx_enable_gta

    integer(c_int) function umt_ctx_dims(ctx, dims) bind(c, name="umt_ctx_dims")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: dims(*)
    end function umt_ctx_dims

!   psi(G, C + nbelem, A)
    integer(c_int) function umt_ctx_get_psi(ctx, psi) bind(c, name="umt_ctx_get_psi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: psi(*)
    end function umt_ctx_get_psi

    integer(c_int) function umt_ctx_set_groups_per_lane(ctx, K) &
        bind(c, name="umt_ctx_set_groups_per_lane")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: K
    end function umt_ctx_set_groups_per_lane

    integer(c_int) function umt_ctx_get_groups_per_lane(ctx, K) &
        bind(c, name="umt_ctx_get_groups_per_lane")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: K
    end function umt_ctx_get_groups_per_lane

    integer(c_int) function umt_ctx_get_angles_per_block(ctx, n, blocks) &
        bind(c, name="umt_ctx_get_angles_per_block")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: n, blocks
    end function umt_ctx_get_angles_per_block

    integer(c_int) function umt_ctx_sweep_n(ctx, n) bind(c, name="umt_ctx_sweep_n")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
    end function umt_ctx_sweep_n

    integer(c_int) function umt_ctx_time_sweeps(ctx, warmup, reps, msSweep, msLevelKernels, &
                                                launches) bind(c, name="umt_ctx_time_sweeps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: warmup, reps
      real(c_double) :: msSweep, msLevelKernels
      integer(c_int) :: launches
    end function umt_ctx_time_sweeps

!   Time-dependent cycles: tau = 1/(c dt); Psi(G, C, A) is Teton's Psi.
    integer(c_int) function umt_ctx_set_tau(ctx, tau) bind(c, name="umt_ctx_set_tau")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tau
    end function umt_ctx_set_tau

    integer(c_int) function umt_ctx_set_psi_prev(ctx, Psi) bind(c, name="umt_ctx_set_psi_prev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: Psi(*)
    end function umt_ctx_set_psi_prev

    integer(c_int) function umt_ctx_get_psi_prev(ctx, Psi) bind(c, name="umt_ctx_get_psi_prev")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: Psi(*)
    end function umt_ctx_get_psi_prev

    integer(c_int) function umt_ctx_save_psi(ctx) bind(c, name="umt_ctx_save_psi")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function umt_ctx_save_psi

!   Batches of uneven domains: problems(nDomains) handles, groups(nDomains).
    integer(c_int) function umt_batch_create(nDomains, problems, groups, precision, batch) &
        bind(c, name="umt_batch_create")
      import :: c_int, c_ptr
      integer(c_int), value :: nDomains, precision
      type(c_ptr) :: problems(*)
      integer(c_int) :: groups(*)
      type(c_ptr) :: batch
    end function umt_batch_create

    subroutine umt_batch_destroy(batch) bind(c, name="umt_batch_destroy")
      import :: c_ptr
      type(c_ptr), value :: batch
    end subroutine umt_batch_destroy

    integer(c_int) function umt_batch_dims(batch, dims) bind(c, name="umt_batch_dims")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: batch
      integer(c_int64_t) :: dims(*)
    end function umt_batch_dims

    integer(c_int) function umt_batch_set_cross_sections(batch, domain, sigt, sigs, q) &
        bind(c, name="umt_batch_set_cross_sections")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: domain
      real(c_double) :: sigt(*), sigs(*), q(*)
    end function umt_batch_set_cross_sections

    integer(c_int) function umt_batch_reset(batch) bind(c, name="umt_batch_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function umt_batch_reset

    integer(c_int) function umt_batch_sweep(batch) bind(c, name="umt_batch_sweep")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function umt_batch_sweep

    integer(c_int) function umt_batch_get_phi(batch, domain, phi) bind(c, name="umt_batch_get_phi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: domain
      real(c_double) :: phi(*)
    end function umt_batch_get_phi

    integer(c_int) function umt_batch_source_iteration(batch, tol, maxIters, iters, residuals, &
                                                       total) bind(c, name="umt_batch_source_iteration")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double), value :: tol
      integer(c_int), value :: maxIters
      integer(c_int) :: iters(*)
      real(c_double) :: residuals(*)
      integer(c_int) :: total
    end function umt_batch_source_iteration

    integer(c_int) function umt_batch_set_graphs(batch, on) bind(c, name="umt_batch_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: on
    end function umt_batch_set_graphs

    integer(c_int) function umt_batch_set_layout(batch, groupsPerLane, anglesPerBlock) &
        bind(c, name="umt_batch_set_layout")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: groupsPerLane, anglesPerBlock
    end function umt_batch_set_layout

    integer(c_int) function umt_batch_time_iterations(batch, warmup, reps, msPerIteration) &
        bind(c, name="umt_batch_time_iterations")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: warmup, reps
      real(c_double) :: msPerIteration
    end function umt_batch_time_iterations

    integer(c_int) function umt_batch_autotune(batch, reps, bestK, bestAnglesPerBlock, &
                                               msPerIteration) bind(c, name="umt_batch_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: reps
      integer(c_int) :: bestK, bestAnglesPerBlock
      real(c_double) :: msPerIteration
    end function umt_batch_autotune

    integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function c_strlen

  end interface

contains

! Message of the calling thread's last failure in the library.
  function umt_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = umt_last_error()
    n = int(c_strlen(p))
    allocate(character(len=n) :: msg)
    if (n > 0) then
      call c_f_pointer(p, chars, [n])
      do i = 1, n
        msg(i:i) = chars(i)
      enddo
    endif
  end function umt_error_message

end module umt_sweep_mod
