!=======================================================================
! Test-only stand-ins for the two modules every reference routine uses
! (same names as src/parameters.f90 and src/globals.f90), sized for the
! C1 point explosion, so the drop-ins can be exercised without the
! reference tree.  TEST SCAFFOLDING.
!=======================================================================
module parameters
  implicit none
  integer, parameter :: neq=4, ndim=2, npas=0, nghost=2
  integer, parameter :: nblocks=64
  integer, parameter :: nx=16, ny=16
  integer, parameter :: nlevs=1, nbroot=4
  real, parameter :: amh=1.66e-24, Rg=8.3145e7, T0=1.e4, amol=0.6
  real, parameter :: cv=1.5, gamma=(cv+1.)/cv
  real, parameter :: xmax=1.0, ymax=1.0, xphys=3.0857E18
  real, parameter :: rsc=xphys/xmax
  real, parameter :: Tempsc=T0*gamma
  real, parameter :: Psc=gamma*amh*Rg*T0
  real, parameter :: vsc2=gamma*T0*Rg/amol
  real, parameter :: tsc=rsc/sqrt(vsc2)
  real, parameter :: cfl=0.4, eta=0.001
  integer, parameter :: neqpas=4+npas
  integer, parameter :: nxmin=1-nghost, nxmax=nx+nghost, nymin=1-nghost, nymax=ny+nghost
end module parameters

module globals
  implicit none
  integer :: nbmin, nbmaxtot, nbmax, nbleafs
  real, dimension(:,:,:,:), allocatable :: u, up, primit
  integer, dimension(:,:), allocatable :: lp
  logical, dimension(:), allocatable :: leaf
  integer, dimension(:,:), allocatable :: icoord
  real, dimension(:), allocatable :: dx, dy
  integer, dimension(:,:), allocatable :: innerbounds
  integer :: ninternal
  integer :: rank, np
end module globals

! The drop-in tstep calls the host's locatebounds when it hands the list
! over; this driver lets the library build the list instead.
subroutine locatebounds(time)
  real, intent(in) :: time
end subroutine locatebounds
