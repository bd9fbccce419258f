!***********************************************************************
!  umt_rz_driver - Fortran host driver for the MI355X RZ UCB sweep.
!
!  Reads a quad mesh in the (r, z) plane, builds the RZ problem from it,
!  reads back the Teton-layout corner geometry (the r-weighted corner
!  Volume for the total it prints) and then
!    mode "si":   steady source iteration on the GPU from Phi = 0
!                 (SetSweep semantics for every direction, scattering
!                 source wtiso (sigs Phi + q) with wtiso = 1/(2 pi));
!    mode "time": `cycles` time-dependent cycles with tau: each is source
!                 iteration with the source STotal + tau Psi(:,:,Angle) of
!                 snac/SweepUCBrz.F90:136 and the cycle's savePsi sweep.
!                 The host holds the starting Psi(G, C, A).
!  Writes Phi(G, C) as raw FP64 to the output file ("time" appends the
!  final Psi(G, C, A)) and prints a one-line summary.
!
!  mesh file (stream, native endianness): int32 npoints, int32 nzones,
!  real64 rz(2, npoints), int32 quads(4, nzones) with 0-based node ids.
!
!  usage: umt_rz_driver mode meshfile npolar nazimuthal groups precision
!                       maxIters tol seed sigt_lo sigt_hi c_lo c_hi q_lo q_hi
!                       tau cycles outfile
!***********************************************************************
program umt_rz_driver

  use iso_c_binding
  use umt_sweep_mod, only: umt_draw_cross_sections, umt_error_message
  use umt_rz_mod

  implicit none

  character(len=256) :: mode, meshfile, outfile
  integer(c_int)     :: npolar, nazimu, G, precision, maxIters, cycles, npoints, nzones
  real(c_double)     :: tol, sigt_lo, sigt_hi, c_lo, c_hi, q_lo, q_hi, tau, residual
  integer(c_int64_t) :: seed, dims(8)
  type(c_ptr)        :: prob, ctx
  integer            :: Z, C, A, rc, u, cyc, ia, ig
  integer(c_int)     :: iters

  real(c_double), allocatable :: rz(:,:), Sigt(:,:), sigs(:,:), q(:,:), Phi(:,:), Psi(:,:,:)
  real(c_double), allocatable :: A_fp(:,:,:), A_ez(:,:,:), RFP(:,:), REZ(:,:), Area(:)
  real(c_double), allocatable :: Volume(:), VolumeZone(:), hist(:)
  integer(c_int), allocatable :: quads(:,:)

  call get_command_argument(1, mode)
  call get_command_argument(2, meshfile)
  call int_arg(3, npolar); call int_arg(4, nazimu); call int_arg(5, G)
  call int_arg(6, precision); call int_arg(7, maxIters)
  call real_arg(8, tol)
  block
    character(len=64) :: s
    call get_command_argument(9, s); read(s, *) seed
  end block
  call real_arg(10, sigt_lo); call real_arg(11, sigt_hi)
  call real_arg(12, c_lo); call real_arg(13, c_hi)
  call real_arg(14, q_lo); call real_arg(15, q_hi)
  call real_arg(16, tau); call int_arg(17, cycles)
  call get_command_argument(18, outfile)

  open(newunit=u, file=trim(meshfile), access="stream", form="unformatted", status="old")
  read(u) npoints, nzones
  allocate(rz(2, npoints), quads(4, nzones))
  read(u) rz, quads
  close(u)

  rc = umt_rz_problem_from_quad_mesh(npoints, rz, nzones, quads, npolar, nazimu, prob)
  call check(rc, "umt_rz_problem_from_quad_mesh")
  call check(umt_rz_problem_dims(prob, dims), "umt_rz_problem_dims")
  Z = int(dims(1)); C = int(dims(2)); A = int(dims(4))

  allocate(A_fp(2,2,C), A_ez(2,2,C), RFP(2,C), REZ(2,C), Area(C), Volume(C), VolumeZone(Z))
  allocate(Sigt(G, Z), sigs(G, Z), q(G, Z), Phi(G, C), hist(max(maxIters, 1)))
  call check(umt_rz_problem_geometry(prob, A_fp, A_ez, RFP, REZ, Area, Volume, VolumeZone), &
             "umt_rz_problem_geometry")
  call check(umt_draw_cross_sections(seed, Z, G, sigt_lo, sigt_hi, c_lo, c_hi, q_lo, q_hi, &
                                     Sigt, sigs, q), "umt_draw_cross_sections")

  call check(umt_rz_ctx_create(prob, G, precision, ctx), "umt_rz_ctx_create")
  call check(umt_rz_ctx_set_sigt(ctx, Sigt), "umt_rz_ctx_set_sigt")
  call check(umt_rz_ctx_set_scattering(ctx, sigs, q), "umt_rz_ctx_set_scattering")

  if (trim(mode) == "si") then
    rc = umt_rz_ctx_source_iteration(ctx, tol, maxIters, iters, residual, hist)
    if (rc /= 0 .and. rc /= 1) call check(rc, "umt_rz_ctx_source_iteration")
    call check(umt_rz_ctx_get_phi(ctx, Phi), "umt_rz_ctx_get_phi")
    write(*, '(a,i6,a,es12.5,a,es22.15)') "rz si iterations=", iters, " residual=", residual, &
         " volumePhi=", sum(matmul(Phi, Volume))
  elseif (trim(mode) == "time") then
    allocate(Psi(G, C, A))
    ! the previous cycle's angular flux held by the host, different per direction
    do ia = 1, A
      do ig = 1, G
        Psi(ig, :, ia) = (1.0d0 + 0.25d0 * mod(ia, 3) + 0.1d0 * ig) / (8.0d0 * atan(1.0d0))
      enddo
    enddo
    call check(umt_rz_ctx_set_tau(ctx, tau), "umt_rz_ctx_set_tau")
    call check(umt_rz_ctx_set_psi_prev(ctx, Psi), "umt_rz_ctx_set_psi_prev")
    do cyc = 1, cycles
      rc = umt_rz_ctx_source_iteration(ctx, tol, maxIters, iters, residual, hist)
      if (rc /= 0 .and. rc /= 1) call check(rc, "umt_rz_ctx_source_iteration")
      call check(umt_rz_ctx_save_psi(ctx), "umt_rz_ctx_save_psi")
      write(*, '(a,i3,a,i6,a,es12.5)') "rz cycle ", cyc, " iterations=", iters, &
           " residual=", residual
    enddo
    call check(umt_rz_ctx_get_phi(ctx, Phi), "umt_rz_ctx_get_phi")
    call check(umt_rz_ctx_get_psi_prev(ctx, Psi), "umt_rz_ctx_get_psi_prev")
  else
    write(*, *) "unknown mode ", trim(mode)
    stop 2
  endif

  open(newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
  write(u) Phi
  if (allocated(Psi)) write(u) Psi
  close(u)
  call umt_rz_ctx_destroy(ctx)
  call umt_rz_problem_destroy(prob)

contains

  subroutine int_arg(k, v)
    integer, intent(in) :: k
    integer(c_int), intent(out) :: v
    character(len=64) :: s
    call get_command_argument(k, s)
    read(s, *) v
  end subroutine int_arg

  subroutine real_arg(k, v)
    integer, intent(in) :: k
    real(c_double), intent(out) :: v
    character(len=64) :: s
    call get_command_argument(k, s)
    read(s, *) v
  end subroutine real_arg

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(*, '(a,a,i5,a,a)') what, " failed with status ", code, ": ", umt_error_message()
      stop 1
    endif
  end subroutine check

end program umt_rz_driver
