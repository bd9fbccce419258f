!***********************************************************************
!  umt_si_driver - Fortran host driver for the MI355X UCB sweep.
!
!  Builds a tiled hex problem, holds every array in the Teton layout
!  (Geometry_mod / AngleSet_mod / GroupSet_mod names and shapes) and then
!    mode "si":     hands the arrays to a device-resident context and runs
!                   steady source iteration on the GPU (rt/rtmainsn.F90
!                   semantics for tau = 0, within-group scattering);
!    mode "gta":    the same with grey-transport acceleration, the context
!                   built from the problem handle (S2 schedule included);
!    mode "gta2":   the same with GTA kernel 2 (transfer matrices);
!    mode "dropin": sweeps one angle through gpu_sweepucbxyz with the
!                   argument list of src/teton/gpu/SweepUCBxyzToGPU.F90:198-240;
!    mode "mesh":   as "si", on a hex mesh read from the file named by the
!                   20th argument (int32 npoints, int32 nzones, real64
!                   coords(3, npoints), int32 connectivity(8, nzones), 0-based
!                   Blueprint / VTK hexahedron order; mesh ingestion);
!    mode "time":   `angle` time-dependent cycles with tau (20th argument):
!                   source iteration with the tau Psi source of
!                   snac/SweepUCBxyz.F90:121, then the savePsi sweep, from a
!                   Psi(G, C, A) the host holds; appends the final Psi;
!    mode "setsweep": one pass of SetSweep's angle loop (snac/SetSweep.F90:
!                   113-170) through SweepUCBxyz(Set, setID, Groups, Angle,
!                   savePsi) of umt_teton_mod, the Teton-shaped entry point,
!                   with incident boundary fluxes PsiB(g, b, a) =
!                   0.01 (1 + mod(g + 3 b + 7 a, 11)) and the problem's cycle
!                   lists; on the hex mesh of the file named by the 20th
!                   argument when one is given; appends the final PsiB(G,
!                   nbelem, A) and Psi(G, C, A) after Phi;
!    mode "batch":  `angle` uneven domains (tiles nx + d - 1, ny, nz and
!                   groups + d - 1 for domain d, cross sections drawn with
!                   seed + d) iterated together in one batch; writes every
!                   domain's Phi in turn.
!  Writes the resulting Phi(G,C) as raw FP64 to the output file and prints a
!  one-line summary.
!
!  usage: umt_si_driver mode nx ny nz npolar nazimuthal groups precision
!                       maxIters tol seed sigt_lo sigt_hi c_lo c_hi q_lo q_hi
!                       angle outfile [tau | meshfile]
!***********************************************************************
program umt_si_driver

  use iso_c_binding
  use umt_sweep_mod
  use umt_teton_mod, only: SetData, HypPlane, Geom, Size, Quad, SweepUCBxyz

  implicit none

  character(len=256) :: arg, mode, outfile, extra
  integer(c_int)     :: nx, ny, nz, npolar, nazimu, G, precision, maxIters, angle
  real(c_double)     :: tol, sigt_lo, sigt_hi, c_lo, c_hi, q_lo, q_hi
  integer(c_int64_t) :: seed
  integer(c_int64_t) :: dims(8)
  type(c_ptr)        :: prob, ctx
  integer            :: Z, C, nb, A, totalPlanes, rc, i, k, u, start
  integer(c_int)     :: iters, bestK, status

  integer(c_int), allocatable :: numCorner(:), cOffSet(:), nCFaces(:), cFP(:,:), cEZ(:,:)
  integer(c_int), allocatable :: CToFace(:,:), zoneOpp(:,:), faceOpp(:,:)
  integer(c_int), allocatable :: nHyperPlanes(:), zonesInPlane(:), nextZ(:,:), nextC(:,:)
  real(c_double), allocatable :: px(:,:), A_fp(:,:,:), A_ez(:,:,:), Volume(:), VolumeZone(:)
  real(c_double), allocatable :: omega(:,:), weight(:)
  real(c_double), allocatable :: Sigt(:,:), sigs(:,:), q(:,:), Phi(:,:), STotal(:,:)
  real(c_double), allocatable :: Psi1(:,:), PsiB(:,:), Psi(:,:), hist(:), Psi3(:,:,:)
  real(c_double)              :: residual, msPerK(4), tau, quadwt
  real(c_double)              :: dummy(1)
  integer(c_int64_t)          :: greyIts, greySweeps
  integer(c_int)              :: idummy(1)
  integer(c_int)              :: ndim, maxcf, maxCorner, zero, one, minusOne, streamId, nHP

  call get_command_argument(1, mode)
  call int_arg(2, nx); call int_arg(3, ny); call int_arg(4, nz)
  call int_arg(5, npolar); call int_arg(6, nazimu); call int_arg(7, G)
  call int_arg(8, precision); call int_arg(9, maxIters)
  call real_arg(10, tol)
  call get_command_argument(11, arg); read(arg, *) seed
  call real_arg(12, sigt_lo); call real_arg(13, sigt_hi)
  call real_arg(14, c_lo); call real_arg(15, c_hi)
  call real_arg(16, q_lo); call real_arg(17, q_hi)
  call int_arg(18, angle)
  call get_command_argument(19, outfile)
  call get_command_argument(20, extra)

  if (trim(mode) == "mesh" .or. (trim(mode) == "setsweep" .and. len_trim(extra) > 0)) then
    block
      integer(c_int) :: npts, nzs
      real(c_double), allocatable :: coords(:,:)
      integer(c_int), allocatable :: conn(:,:)
      open(newunit=u, file=trim(extra), access="stream", form="unformatted", status="old")
      read(u) npts, nzs
      allocate(coords(3, npts), conn(8, nzs))
      read(u) coords, conn
      close(u)
      rc = umt_problem_from_hex_mesh(npts, coords, nzs, conn, npolar, nazimu, 1, prob)
      call check(rc, "umt_problem_from_hex_mesh")
    end block
  else
    rc = umt_problem_tiled(nx, ny, nz, npolar, nazimu, 1, prob)
    call check(rc, "umt_problem_tiled")
  endif
  rc = umt_problem_dims(prob, dims)
  Z = int(dims(1)); C = int(dims(2)); nb = int(dims(3)); A = int(dims(4))
  totalPlanes = int(dims(7))

  allocate(numCorner(Z), cOffSet(Z), nCFaces(C), cFP(3,C), cEZ(3,C), CToFace(3,C))
  allocate(zoneOpp(6,Z), faceOpp(6,Z), px(3,C), A_fp(3,3,C), A_ez(3,3,C), Volume(C))
  allocate(VolumeZone(Z), omega(3,A), weight(A), nHyperPlanes(A), zonesInPlane(totalPlanes))
  allocate(nextZ(Z,A), nextC(C,A), Sigt(G,Z), sigs(G,Z), q(G,Z), Phi(G,C))

  rc = umt_problem_mesh(prob, numCorner, cOffSet, nCFaces, cFP, cEZ, CToFace, zoneOpp, faceOpp, px)
  call check(rc, "umt_problem_mesh")
  rc = umt_problem_geometry(prob, A_fp, A_ez, Volume, VolumeZone)
  call check(rc, "umt_problem_geometry")
  rc = umt_problem_quadrature(prob, omega, weight)
  call check(rc, "umt_problem_quadrature")
  rc = umt_problem_schedule(prob, nHyperPlanes, zonesInPlane, nextZ, nextC)
  call check(rc, "umt_problem_schedule")
  rc = umt_draw_cross_sections(seed, Z, G, sigt_lo, sigt_hi, c_lo, c_hi, q_lo, q_hi, Sigt, sigs, q)
  call check(rc, "umt_draw_cross_sections")

  if (trim(mode) == "si" .or. trim(mode) == "mesh") then

    rc = umt_ctx_create(Z, C, nb, G, A, precision, numCorner, cOffSet, nCFaces, cFP, cEZ, &
                        A_fp, A_ez, Volume, VolumeZone, omega, weight, nHyperPlanes,       &
                        zonesInPlane, nextZ, ctx)
    call check(rc, "umt_ctx_create")
    call check(umt_ctx_set_sigt(ctx, Sigt), "umt_ctx_set_sigt")
    call check(umt_ctx_set_scattering(ctx, sigs, q), "umt_ctx_set_scattering")
    call check(umt_ctx_autotune(ctx, 1, bestK, msPerK), "umt_ctx_autotune")
    allocate(hist(max(maxIters, 1)))
    rc = umt_ctx_source_iteration(ctx, tol, maxIters, iters, residual, hist)
    if (rc /= 0 .and. rc /= 1) call check(rc, "umt_ctx_source_iteration")
    call check(umt_ctx_get_phi(ctx, Phi), "umt_ctx_get_phi")
    call umt_ctx_destroy(ctx)
    write(*, '(a,i6,a,es12.5,a,i2,a,es22.15)') "si iterations=", iters, " residual=", residual, &
         " groupsPerLane=", bestK, " sumPhi=", sum(Phi)

  elseif (trim(mode) == "gta" .or. trim(mode) == "gta2") then

    ! Context from the problem (it brings the S2 grey-acceleration schedule),
    ! GTA-accelerated source iteration with graph replay.
    rc = umt_ctx_create_from_problem(prob, G, precision, ctx)
    call check(rc, "umt_ctx_create_from_problem")
    call check(umt_ctx_set_sigt(ctx, Sigt), "umt_ctx_set_sigt")
    call check(umt_ctx_set_scattering(ctx, sigs, q), "umt_ctx_set_scattering")
    call check(umt_ctx_set_acceleration(ctx, 1), "umt_ctx_set_acceleration")
    if (trim(mode) == "gta2") call check(umt_ctx_set_gta_solver(ctx, 2), "umt_ctx_set_gta_solver")
    call check(umt_ctx_set_graphs(ctx, 1), "umt_ctx_set_graphs")
    allocate(hist(max(maxIters, 1)))
    rc = umt_ctx_source_iteration(ctx, tol, maxIters, iters, residual, hist)
    if (rc /= 0 .and. rc /= 1) call check(rc, "umt_ctx_source_iteration")
    call check(umt_ctx_gta_stats(ctx, greyIts, greySweeps), "umt_ctx_gta_stats")
    call check(umt_ctx_get_phi(ctx, Phi), "umt_ctx_get_phi")
    call umt_ctx_destroy(ctx)
    write(*, '(a,i6,a,es12.5,a,i8,a,es22.15)') "gta iterations=", iters, " residual=", residual, &
         " greyIterations=", greyIts, " sumPhi=", sum(Phi)

  elseif (trim(mode) == "dropin") then

    ! Source for one sweep: STotal = (sigs * phi0 + q) / (4 pi) with phi0 = 1.
    allocate(STotal(G,C), Psi1(G,C+nb), PsiB(G,nb), Psi(G,C))
    do i = 1, C
      STotal(:,i) = (sigs(:, (i-1)/8 + 1) + q(:, (i-1)/8 + 1)) / (16.0d0*atan(1.0d0))
    enddo
    Psi1 = 0.0d0; PsiB = 0.0d0; Psi = 0.0d0; Phi = 0.0d0
    start = 0
    do i = 1, angle - 1
      start = start + nHyperPlanes(i)
    enddo
    nHP = nHyperPlanes(angle)
    tau = 0.0d0; quadwt = weight(angle)
    ndim = 3; maxcf = 3; maxCorner = 8; zero = 0; one = 1; minusOne = -1; streamId = 0
    call gpu_sweepucbxyz(angle, nHP, zonesInPlane(start+1), nextZ(1,angle), nextC(1,angle), &
                         STotal, tau, Psi, G, Volume, Sigt, nCFaces, ndim, maxcf, C,          &
                         A_fp, omega(1,angle), cFP, Psi1, nb, A_ez, cEZ, A, quadwt, Phi,      &
                         PsiB, maxCorner, one, streamId, one, zero, zero, zero, dummy,        &
                         idummy, zero, zero, dummy, minusOne, numCorner, cOffSet)
    call gpu_streamsynchronize(streamId)
    status = gpu_sweepucbxyz_status()
    if (status /= 0) call check(status, "gpu_sweepucbxyz")
    write(*, '(a,i4,a,es22.15)') "dropin angle=", angle, " sumPhi=", sum(Phi)

  elseif (trim(mode) == "setsweep") then

    ! Teton-shaped data (Geom, Quad with one angle and group set, Set) from
    ! the host arrays, then SetSweep's loop: Phi = 0, SweepUCBxyz per angle.
    block
      type(SetData), target :: Set
      integer(c_int) :: totalCycles
      integer        :: ia, ib, ig, ih, ncl
      allocate(Geom% numCorner(Z), Geom% cOffSet(Z), Geom% nCFacesArray(C), Geom% cFP(3,C), &
               Geom% cEZ(3,C), Geom% Volume(C), Geom% A_fp(3,3,C), Geom% A_ez(3,3,C))
      Geom% numCorner = numCorner; Geom% cOffSet = cOffSet; Geom% nCFacesArray = nCFaces
      Geom% cFP = cFP; Geom% cEZ = cEZ; Geom% Volume = Volume; Geom% A_fp = A_fp
      Geom% A_ez = A_ez
      Size% tau = 0.0d0
      allocate(Quad% AngSet(1), Quad% GrpSet(1))
      associate (ASet => Quad% AngSet(1), GSet => Quad% GrpSet(1))
        ASet% NumAngles = A
        allocate(ASet% omega(3,A), ASet% weight(A), ASet% nextZ(Z,A), ASet% nextC(C,A), &
                 ASet% nHyperPlanes(A), ASet% numCycles(A), ASet% cycleOffSet(A),         &
                 ASet% HypPlanePtr(A))
        ASet% omega = omega; ASet% weight = weight; ASet% nextZ = nextZ; ASet% nextC = nextC
        ASet% nHyperPlanes = nHyperPlanes
        start = 0
        do ia = 1, A
          allocate(ASet% HypPlanePtr(ia)% zonesInPlane(nHyperPlanes(ia)))
          do ih = 1, nHyperPlanes(ia)
            ASet% HypPlanePtr(ia)% zonesInPlane(ih) = zonesInPlane(start + ih)
          enddo
          start = start + nHyperPlanes(ia)
        enddo
        ! cycle lists of the schedule (rt/rtorder.F90:62-90 offsets)
        rc = umt_problem_cycles(prob, totalCycles, c_loc(ASet% numCycles), c_null_ptr)
        call check(rc, "umt_problem_cycles")
        ncl = max(int(totalCycles), 1)
        allocate(ASet% cycleList(ncl))
        ASet% cycleList = 0
        rc = umt_problem_cycles(prob, totalCycles, c_loc(ASet% numCycles), c_loc(ASet% cycleList))
        call check(rc, "umt_problem_cycles")
        ASet% cycleOffSet(1) = 0
        do ia = 2, A
          ASet% cycleOffSet(ia) = ASet% cycleOffSet(ia-1) + ASet% numCycles(ia-1)
        enddo
        allocate(GSet% STotal(G,C), GSet% Sigt(G,Z))
        do i = 1, C
          GSet% STotal(:,i) = (sigs(:, (i-1)/8 + 1) + q(:, (i-1)/8 + 1)) / (16.0d0*atan(1.0d0))
        enddo
        GSet% Sigt = Sigt
        Set% Groups = G; Set% nCorner = C; Set% nbelem = nb; Set% NumAngles = A
        allocate(Set% Psi(G,C,A), Set% Psi1(G,C+nb), Set% PsiB(G,nb,A), Set% Phi(G,C), &
                 Set% cyclePsi(G,ncl))
        Set% Psi = 0.0d0; Set% Psi1 = 0.0d0; Set% Phi = 0.0d0; Set% cyclePsi = 0.0d0
        do ia = 1, A
          do ib = 1, nb
            do ig = 1, G
              Set% PsiB(ig,ib,ia) = 0.01d0 * (1 + mod(ig + 3*ib + 7*ia, 11))
            enddo
          enddo
        enddo
        do ia = 1, A
          call SweepUCBxyz(Set, 1, G, ia, logical(.true., kind=1))
        enddo
        Phi = Set% Phi
        allocate(Psi3(G, nb*A + C*A, 1))
        Psi3(:,:,1) = reshape([reshape(Set% PsiB, [G*nb*A]), reshape(Set% Psi, [G*C*A])], &
                              [G, nb*A + C*A])
      end associate
      write(*, '(a,i4,a,es22.15)') "setsweep angles=", A, " sumPhi=", sum(Phi)
    end block

  elseif (trim(mode) == "time") then

    ! The previous cycle's angular flux held by the host, different per angle.
    read(extra, *) tau
    allocate(Psi3(G, C, A), hist(max(maxIters, 1)))
    do i = 1, A
      do k = 1, G
        Psi3(k, :, i) = (1.0d0 + 0.25d0 * mod(i, 3) + 0.1d0 * k) / (16.0d0 * atan(1.0d0))
      enddo
    enddo
    rc = umt_ctx_create_from_problem(prob, G, precision, ctx)
    call check(rc, "umt_ctx_create_from_problem")
    call check(umt_ctx_set_sigt(ctx, Sigt), "umt_ctx_set_sigt")
    call check(umt_ctx_set_scattering(ctx, sigs, q), "umt_ctx_set_scattering")
    call check(umt_ctx_set_tau(ctx, tau), "umt_ctx_set_tau")
    call check(umt_ctx_set_psi_prev(ctx, Psi3), "umt_ctx_set_psi_prev")
    do i = 1, angle
      rc = umt_ctx_source_iteration(ctx, tol, maxIters, iters, residual, hist)
      if (rc /= 0 .and. rc /= 1) call check(rc, "umt_ctx_source_iteration")
      call check(umt_ctx_save_psi(ctx), "umt_ctx_save_psi")
      write(*, '(a,i3,a,i6,a,es12.5)') "time cycle ", i, " iterations=", iters, &
           " residual=", residual
    enddo
    call check(umt_ctx_get_phi(ctx, Phi), "umt_ctx_get_phi")
    call check(umt_ctx_get_psi_prev(ctx, Psi3), "umt_ctx_get_psi_prev")
    call umt_ctx_destroy(ctx)

  elseif (trim(mode) == "batch") then

    block
      type(c_ptr), allocatable :: probs(:)
      integer(c_int), allocatable :: gs(:), its(:)
      real(c_double), allocatable :: res(:), sT(:,:), sS(:,:), sQ(:,:), ph(:,:)
      type(c_ptr) :: batch
      integer(c_int) :: d, total, bestK, bestApb
      integer(c_int64_t) :: dd(8)
      real(c_double) :: ms
      allocate(probs(angle), gs(angle), its(angle), res(angle))
      do d = 1, angle
        call check(umt_problem_tiled(nx + d - 1, ny, nz, npolar, nazimu, 1, probs(d)), &
                   "umt_problem_tiled")
        gs(d) = G + d - 1
      enddo
      call check(umt_batch_create(angle, probs, gs, precision, batch), "umt_batch_create")
      do d = 1, angle
        call check(umt_problem_dims(probs(d), dd), "umt_problem_dims")
        allocate(sT(gs(d), dd(1)), sS(gs(d), dd(1)), sQ(gs(d), dd(1)))
        call check(umt_draw_cross_sections(seed + d, int(dd(1), c_int), gs(d), sigt_lo, sigt_hi, &
                                           c_lo, c_hi, q_lo, q_hi, sT, sS, sQ),            &
                   "umt_draw_cross_sections")
        call check(umt_batch_set_cross_sections(batch, d - 1, sT, sS, sQ), &
                   "umt_batch_set_cross_sections")
        deallocate(sT, sS, sQ)
      enddo
      call check(umt_batch_autotune(batch, 1, bestK, bestApb, ms), "umt_batch_autotune")
      rc = umt_batch_source_iteration(batch, tol, maxIters, its, res, total)
      if (rc /= 0 .and. rc /= 1) call check(rc, "umt_batch_source_iteration")
      write(*, '(a,i3,a,i6,a,i2,a,i4)') "batch domains=", angle, " iterations=", total, &
           " groupsPerLane=", bestK, " anglesPerBlock=", bestApb
      open(newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
      do d = 1, angle
        call check(umt_problem_dims(probs(d), dd), "umt_problem_dims")
        allocate(ph(gs(d), dd(2)))
        call check(umt_batch_get_phi(batch, d - 1, ph), "umt_batch_get_phi")
        write(u) ph
        deallocate(ph)
      enddo
      close(u)
      call umt_batch_destroy(batch)
      do d = 1, angle
        call umt_problem_destroy(probs(d))
      enddo
    end block
    call umt_problem_destroy(prob)
    stop

  else
    write(*, *) "unknown mode ", trim(mode)
    stop 2
  endif

  open(newunit=u, file=trim(outfile), access="stream", form="unformatted", status="replace")
  write(u) Phi
  if (allocated(Psi3)) write(u) Psi3
  close(u)
  call umt_problem_destroy(prob)

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

end program umt_si_driver
