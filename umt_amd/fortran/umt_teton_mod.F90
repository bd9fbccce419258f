!***********************************************************************
!  umt_teton_mod - Teton's sweep entry point SweepUCBxyz(Set, setID,
!                  Groups, Angle, savePsi) (snac/SweepUCBxyz.F90:11) over the
!                  MI355X sweep, for Fortran hosts that hold their data in
!                  Teton's shapes but not in Teton's modules.
!
!  The derived types carry the components of Teton's SetData
!  (mods/SetData_mod.F90), AngleSet and HypPlane (mods/AngleSet_mod.F90:
!  60-102), GroupSet (mods/GroupSet_mod.F90), Geometry (mods/Geometry_mod.F90:
!  23-48) and MeshSize (mods/Size_mod.F90) that the sweep reads, under the
!  same names, and the accessors keep Teton's names
!  (getAngleSetFromSetID, getGroupSetFromSetID: mods/QuadratureList_mod.F90;
!  getNumberOfHyperPlanes: mods/AngleSet_mod.F90:139). The body of
!  SweepUCBxyz is umt_sweepucbxyz_body.inc, which teton/SweepUCBxyz.F90 also
!  includes, so a Teton build and this module run the same statements.
!  A host fills Geom, Size and Quad (one AngleSet and GroupSet per set) and
!  calls SweepUCBxyz for each angle as SetSweep does (snac/SetSweep.F90:
!  113-170).
!***********************************************************************
module umt_teton_mod

  use iso_c_binding
  use umt_sweep_mod, only: gpu_sweepucbxyz, gpu_streamsynchronize, gpu_sweepucbxyz_status, &
                           umt_error_message

  implicit none

  private

  public :: SetData, AngleSet, HypPlane, GroupSet, Geometry, MeshSize, QuadratureList
  public :: Geom, Size, Quad
  public :: getAngleSetFromSetID, getGroupSetFromSetID, getNumberOfHyperPlanes
  public :: SweepUCBxyz

  type :: HypPlane
    integer(c_int), pointer, contiguous :: zonesInPlane(:) => null()
  end type HypPlane

  type :: AngleSet
    integer                             :: NumAngles = 0
    real(c_double), pointer, contiguous :: omega(:,:) => null()     ! (3, NumAngles)
    real(c_double), pointer, contiguous :: weight(:) => null()
    integer(c_int), pointer, contiguous :: nextZ(:,:) => null()     ! (nZones, NumAngles)
    integer(c_int), pointer, contiguous :: nextC(:,:) => null()     ! (nCorner, NumAngles)
    integer(c_int), pointer, contiguous :: nHyperPlanes(:) => null()
    integer(c_int), pointer, contiguous :: numCycles(:) => null()
    integer(c_int), pointer, contiguous :: cycleOffSet(:) => null()
    integer(c_int), pointer, contiguous :: cycleList(:) => null()
    type(HypPlane), pointer             :: HypPlanePtr(:) => null()
  end type AngleSet

  type :: GroupSet
    real(c_double), pointer, contiguous :: STotal(:,:) => null()    ! (Groups, nCorner)
    real(c_double), pointer, contiguous :: Sigt(:,:) => null()      ! (Groups, nZones)
  end type GroupSet

  type :: SetData
    integer                             :: Groups = 0, nCorner = 0, nbelem = 0, NumAngles = 0
    real(c_double), pointer, contiguous :: Psi(:,:,:) => null()     ! (Groups, nCorner, NumAngles)
    real(c_double), pointer, contiguous :: Psi1(:,:) => null()      ! (Groups, nCorner + nbelem)
    real(c_double), pointer, contiguous :: PsiB(:,:,:) => null()    ! (Groups, nbelem, NumAngles)
    real(c_double), pointer, contiguous :: Phi(:,:) => null()       ! (Groups, nCorner)
    real(c_double), pointer, contiguous :: cyclePsi(:,:) => null()  ! (Groups, totalCycles)
  end type SetData

  type :: Geometry
    integer(c_int), pointer, contiguous :: numCorner(:) => null(), cOffSet(:) => null()
    integer(c_int), pointer, contiguous :: nCFacesArray(:) => null()
    integer(c_int), pointer, contiguous :: cFP(:,:) => null(), cEZ(:,:) => null()
    real(c_double), pointer, contiguous :: Volume(:) => null()
    real(c_double), pointer, contiguous :: A_fp(:,:,:) => null(), A_ez(:,:,:) => null()
  end type Geometry

  type :: MeshSize
    integer(c_int) :: ndim = 3, maxcf = 3, maxCorner = 8
    real(c_double) :: tau = 0.0d0
  end type MeshSize

  type :: QuadratureList
    type(AngleSet), pointer :: AngSet(:) => null()   ! by setID
    type(GroupSet), pointer :: GrpSet(:) => null()
  end type QuadratureList

  type(Geometry),       target, save :: Geom
  type(MeshSize),       target, save :: Size
  type(QuadratureList), target, save :: Quad

contains

  function getAngleSetFromSetID(self, setID) result(ASet)
    type(QuadratureList), intent(in) :: self
    integer,              intent(in) :: setID
    type(AngleSet),       pointer    :: ASet
    ASet => self% AngSet(setID)
  end function getAngleSetFromSetID

  function getGroupSetFromSetID(self, setID) result(GSet)
    type(QuadratureList), intent(in) :: self
    integer,              intent(in) :: setID
    type(GroupSet),       pointer    :: GSet
    GSet => self% GrpSet(setID)
  end function getGroupSetFromSetID

  integer(c_int) function getNumberOfHyperPlanes(ASet, angle)
    type(AngleSet), intent(in) :: ASet
    integer,        intent(in) :: angle
    getNumberOfHyperPlanes = ASet% nHyperPlanes(angle)
  end function getNumberOfHyperPlanes

  subroutine SweepUCBxyz(Set, setID, Groups, Angle, savePsi)
#include "umt_sweepucbxyz_decl.inc"
#include "umt_sweepucbxyz_body.inc"
  end subroutine SweepUCBxyz

end module umt_teton_mod
