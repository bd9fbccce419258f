!***********************************************************************
!  umt_rz_mod - Fortran host binding of the MI355X 2D cylindrical (RZ)
!               UCB sweep (libumt_amd.so, C-ABI in include/umt_rz.h).
!
!  The reference sweeps RZ directions one at a time with SweepUCBrz
!  (src/teton/snac/SweepUCBrz.F90:11), called from SetSweep for every
!  direction that is not a finishing direction; its GPU form is OpenMP
!  offload with no C entry point. A Fortran host reaches the device-resident
!  RZ context through these interfaces instead: it hands over its quad mesh
!  (points and four node ids per zone), reads back the Teton-layout corner
!  arrays (Geometry_mod names and shapes for ndim = 2) and sweeps, iterates
!  or steps in time on the GPU. Arrays are column-major with groups fastest,
!  as Teton holds them; handles are type(c_ptr).
!***********************************************************************
module umt_rz_mod

  use iso_c_binding

  implicit none

  private

  public :: umt_rz_problem_structured, umt_rz_problem_from_quad_mesh, umt_rz_problem_destroy, &
            umt_rz_problem_dims, umt_rz_problem_mesh, umt_rz_problem_points,                  &
            umt_rz_problem_geometry, umt_rz_problem_quadrature, umt_rz_problem_schedule,      &
            umt_rz_ctx_create, umt_rz_ctx_destroy, umt_rz_ctx_dims, umt_rz_ctx_set_sigt,      &
            umt_rz_ctx_set_source, umt_rz_ctx_set_scattering, umt_rz_ctx_sweep,               &
            umt_rz_ctx_get_phi, umt_rz_ctx_set_phi, umt_rz_ctx_get_psi,                       &
            umt_rz_ctx_source_iteration, umt_rz_ctx_set_tau, umt_rz_ctx_set_psi_prev,         &
            umt_rz_ctx_get_psi_prev, umt_rz_ctx_save_psi, umt_rz_ctx_set_groups_per_lane,     &
            umt_rz_ctx_set_graphs, umt_rz_ctx_time_sweeps, umt_rz_ctx_autotune,   &
            umt_rz_ctx_set_tiles

  interface

    integer(c_int) function umt_rz_problem_structured(nr, nz, r0, r1, z0, z1, jitter, seed, &
                                                      npolar, nazimuthal, prob)             &
        bind(c, name="umt_rz_problem_structured")
      import :: c_int, c_double, c_int64_t, c_ptr
      integer(c_int), value     :: nr, nz, npolar, nazimuthal
      real(c_double), value     :: r0, r1, z0, z1, jitter
      integer(c_int64_t), value :: seed
      type(c_ptr)               :: prob
    end function

!   rz(2, npoints) as (r, z); quads(4, nzones) 0-based node ids, either orientation
    integer(c_int) function umt_rz_problem_from_quad_mesh(npoints, rz, nzones, quads, npolar, &
                                                          nazimuthal, prob)                   &
        bind(c, name="umt_rz_problem_from_quad_mesh")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: npoints, nzones, npolar, nazimuthal
      real(c_double)        :: rz(*)
      integer(c_int)        :: quads(*)
      type(c_ptr)           :: prob
    end function

    subroutine umt_rz_problem_destroy(prob) bind(c, name="umt_rz_problem_destroy")
      import :: c_ptr
      type(c_ptr), value :: prob
    end subroutine

!   dims: nzones, ncorner, nbelem, nangles, npoints, total hyperplanes, xi levels, sweep levels
    integer(c_int) function umt_rz_problem_dims(prob, dims) bind(c, name="umt_rz_problem_dims")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: prob
      integer(c_int64_t) :: dims(*)
    end function

!   numCorner(Z), cOffSet(Z), cFP(2,C), cEZ(2,C), CToFace(2,C), zoneOpp(4,Z), faceOpp(4,Z), px(2,C)
    integer(c_int) function umt_rz_problem_mesh(prob, numCorner, cOffSet, cFP, cEZ, CToFace, &
                                                zoneOpp, faceOpp, px)                        &
        bind(c, name="umt_rz_problem_mesh")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: prob
      integer(c_int)     :: numCorner(*), cOffSet(*), cFP(*), cEZ(*), CToFace(*)
      integer(c_int)     :: zoneOpp(*), faceOpp(*)
      real(c_double)     :: px(*)
    end function

    integer(c_int) function umt_rz_problem_points(prob, rz, quads) &
        bind(c, name="umt_rz_problem_points")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: prob
      real(c_double)     :: rz(*)
      integer(c_int)     :: quads(*)
    end function

!   A_fp(2,2,C), A_ez(2,2,C), RadiusFP(2,C), RadiusEZ(2,C), Area(C), Volume(C), VolumeZone(Z)
    integer(c_int) function umt_rz_problem_geometry(prob, A_fp, A_ez, RadiusFP, RadiusEZ, Area, &
                                                    Volume, VolumeZone)                         &
        bind(c, name="umt_rz_problem_geometry")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: prob
      real(c_double)     :: A_fp(*), A_ez(*), RadiusFP(*), RadiusEZ(*), Area(*), Volume(*)
      real(c_double)     :: VolumeZone(*)
    end function

    integer(c_int) function umt_rz_problem_quadrature(prob, omega, weight, starting, finishing, &
                                                      angDerivFac, quadTauW1, quadTauW2)        &
        bind(c, name="umt_rz_problem_quadrature")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: prob
      real(c_double)     :: omega(*), weight(*), angDerivFac(*), quadTauW1(*), quadTauW2(*)
      integer(c_int)     :: starting(*), finishing(*)
    end function

    integer(c_int) function umt_rz_problem_schedule(prob, nHyperPlanes, zonesInPlane, nextZ, &
                                                    nextC) bind(c, name="umt_rz_problem_schedule")
      import :: c_int, c_ptr
      type(c_ptr), value :: prob
      integer(c_int)     :: nHyperPlanes(*), zonesInPlane(*), nextZ(*), nextC(*)
    end function

    integer(c_int) function umt_rz_ctx_create(prob, groups, precision, ctx) &
        bind(c, name="umt_rz_ctx_create")
      import :: c_int, c_ptr
      type(c_ptr), value    :: prob
      integer(c_int), value :: groups, precision
      type(c_ptr)           :: ctx
    end function

    subroutine umt_rz_ctx_destroy(ctx) bind(c, name="umt_rz_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function umt_rz_ctx_dims(ctx, dims) bind(c, name="umt_rz_ctx_dims")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: dims(*)
    end function

    integer(c_int) function umt_rz_ctx_set_sigt(ctx, sigt) bind(c, name="umt_rz_ctx_set_sigt")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: sigt(*)
    end function

    integer(c_int) function umt_rz_ctx_set_source(ctx, STotal) bind(c, name="umt_rz_ctx_set_source")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: STotal(*)
    end function

    integer(c_int) function umt_rz_ctx_set_scattering(ctx, sigs, q) &
        bind(c, name="umt_rz_ctx_set_scattering")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: sigs(*), q(*)
    end function

    integer(c_int) function umt_rz_ctx_sweep(ctx) bind(c, name="umt_rz_ctx_sweep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function umt_rz_ctx_get_phi(ctx, phi) bind(c, name="umt_rz_ctx_get_phi")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: phi(*)
    end function

    integer(c_int) function umt_rz_ctx_set_phi(ctx, phi) bind(c, name="umt_rz_ctx_set_phi")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: phi(*)
    end function

    integer(c_int) function umt_rz_ctx_get_psi(ctx, psi) bind(c, name="umt_rz_ctx_get_psi")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: psi(*)
    end function

    integer(c_int) function umt_rz_ctx_source_iteration(ctx, tol, maxIters, iters, residual, &
                                                        history)                             &
        bind(c, name="umt_rz_ctx_source_iteration")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value    :: ctx
      real(c_double), value :: tol
      integer(c_int), value :: maxIters
      integer(c_int)        :: iters
      real(c_double)        :: residual, history(*)
    end function

    integer(c_int) function umt_rz_ctx_set_tau(ctx, tau) bind(c, name="umt_rz_ctx_set_tau")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value    :: ctx
      real(c_double), value :: tau
    end function

!   Psi(G, C, A), the previous cycle's angular flux
    integer(c_int) function umt_rz_ctx_set_psi_prev(ctx, Psi) bind(c, name="umt_rz_ctx_set_psi_prev")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: Psi(*)
    end function

    integer(c_int) function umt_rz_ctx_get_psi_prev(ctx, Psi) bind(c, name="umt_rz_ctx_get_psi_prev")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double)     :: Psi(*)
    end function

    integer(c_int) function umt_rz_ctx_save_psi(ctx) bind(c, name="umt_rz_ctx_save_psi")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function umt_rz_ctx_set_groups_per_lane(ctx, K) &
        bind(c, name="umt_rz_ctx_set_groups_per_lane")
      import :: c_int, c_ptr
      type(c_ptr), value    :: ctx
      integer(c_int), value :: K
    end function

!   msPerK(3): mean ms per sweep for groups per lane 1, 2, 4 (-1 when unsupported)
    integer(c_int) function umt_rz_ctx_autotune(ctx, reps, bestK, msPerK) &
        bind(c, name="umt_rz_ctx_autotune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value    :: ctx
      integer(c_int), value :: reps
      integer(c_int)        :: bestK
      real(c_double)        :: msPerK(*)
    end function

    integer(c_int) function umt_rz_ctx_set_tiles(ctx, on) bind(c, name="umt_rz_ctx_set_tiles")
      import :: c_int, c_ptr
      type(c_ptr), value    :: ctx
      integer(c_int), value :: on
    end function

    integer(c_int) function umt_rz_ctx_set_graphs(ctx, on) bind(c, name="umt_rz_ctx_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value    :: ctx
      integer(c_int), value :: on
    end function

    integer(c_int) function umt_rz_ctx_time_sweeps(ctx, warmup, reps, msSweep, msLevelKernels) &
        bind(c, name="umt_rz_ctx_time_sweeps")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value    :: ctx
      integer(c_int), value :: warmup, reps
      real(c_double)        :: msSweep, msLevelKernels
    end function

  end interface

end module umt_rz_mod
