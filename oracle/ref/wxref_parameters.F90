!=======================================================================
! Problem-directory parameters module for building the reference's
! hot-path sources as a test oracle (the reference's own plugin
! mechanism: a problem directory shadows src/parameters.f90, as
! JET/parameters.f90 does).  Values come from -D switches set by
! oracle/ref/Makefile.  TEST INFRASTRUCTURE ONLY.
!=======================================================================
module parameters
  implicit none
#ifdef MPIP
  include "mpif.h"
#endif
  character (len=128),parameter :: outputpath='./'
  integer, parameter :: neq=WXR_NEQ, ndim=2
  integer, parameter :: npas=WXR_NEQ-4
  integer, parameter :: nghost=2
  integer, parameter :: nblocks=WXR_NBLOCKS
  integer, parameter :: nx=16, ny=16
  integer, parameter :: nlevs=WXR_NLEVS
  integer, parameter :: nbroot=1
#ifdef MPIP
  integer, parameter :: nload=50
#endif
  real, parameter :: pc=3.0857E18, AU=1.496e13
  real, parameter :: clight=2.99E10
  real, parameter :: Msun=1.99E33, Rsun=6.955e10
  real, parameter :: amh=1.66e-24,Kb=1.38e-16,Rg=8.3145e7
  real, parameter :: yr=3.156E7
  real, parameter :: pi=acos(-1.)
  real, parameter :: amol=WXR_AMOL
  real, parameter :: cv=1.5, gamma=(cv+1.)/cv
  real, parameter :: xmax=WXR_XMAX, ymax=WXR_YMAX
  real, parameter :: xphys=WXR_XPHYS
  real, parameter :: T0=1.e4
  real, parameter :: rsc=xphys/xmax, rhosc=amh*amol
  real, parameter :: Tempsc=T0*gamma
  real, parameter :: Psc=gamma*amh*Rg*T0
  real, parameter :: vsc2=gamma*T0*Rg/amol
  real, parameter :: tsc=rsc/sqrt(vsc2)
  real, parameter :: tmax=1.e30, dtprint=1.e30
  real, parameter :: cfl=WXR_CFL
  real, parameter :: eta=0.001
  integer, parameter :: iwarm=0
  integer, parameter :: itprint0=0
  integer, parameter :: neqpas=4+npas
  integer, parameter :: master=0
  integer, parameter :: nxmin=1-nghost, nxmax=nx+nghost,               &
                        nymin=1-nghost, nymax=ny+nghost
#ifndef MPIP
  integer, parameter :: nload=1
#endif
#ifdef MPIP
#ifdef DOUBLEP
  integer, parameter :: mpi_real_kind=mpi_real8
#else
  integer, parameter :: mpi_real_kind=mpi_real4
#endif
#endif
end module parameters
