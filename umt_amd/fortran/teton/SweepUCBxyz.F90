!***********************************************************************
!  SweepUCBxyz for a Teton build: compiled in place of
!  src/teton/snac/SweepUCBxyz.F90 (same name and argument list,
!  snac/SweepUCBxyz.F90:11-30), it sweeps the angle on the MI355X through
!  gpu_sweepucbxyz, so Teton's CPU dispatch (rt/ControlSweep.F90:55-73 ->
!  snac/SetSweep.F90:152) reaches the HIP sweep. The statements are
!  umt_sweepucbxyz_body.inc, the same ones umt_teton_mod.F90 runs in this
!  repository's tests; here they use Teton's own modules. Link with
!  libumt_amd.so and compile with -I pointing at umt_amd/fortran (the .inc
!  files and the umt_sweep_mod module).
!***********************************************************************
   subroutine SweepUCBxyz(Set, setID, Groups, Angle, savePsi)

   use iso_c_binding
   use kind_mod
   use Size_mod
   use Geometry_mod
   use QuadratureList_mod
   use SetData_mod
   use AngleSet_mod
   use GroupSet_mod
   use umt_sweep_mod, only: gpu_sweepucbxyz, gpu_streamsynchronize, gpu_sweepucbxyz_status, &
                            umt_error_message

   implicit none

#include "umt_sweepucbxyz_decl.inc"
#include "umt_sweepucbxyz_body.inc"

   end subroutine SweepUCBxyz
