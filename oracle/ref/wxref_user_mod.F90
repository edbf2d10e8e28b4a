!=======================================================================
! Problem-directory user module for the test-oracle build (shadows
! src/user_mod.f90 the way JET/user_mod.f90 does).  impose_user_bc
! imposes a top-hat jet source whose parameters are set from C through
! wxr_set_jet.  TEST INFRASTRUCTURE ONLY.
!=======================================================================
module user_mod
  implicit none
  real, save :: RJ=0., LJ=0., posJx=0., posJy=0., denJ=0., TempJ=0.
  real, save :: vJ0=0., dvJ=0., tauJ=1., yJ=0., efJ=0.
contains
  subroutine init_user_mod()
  end subroutine init_user_mod

  subroutine initial_conditions(u,nb,time)
    use parameters, only : neq, nxmin, nxmax, nymin, nymax
    implicit none
    real,    intent(out) :: u(neq,nxmin:nxmax,nymin:nymax)
    integer, intent(in)  :: nb
    real,    intent(in)  :: time
    u(:,:,:)=0.
  end subroutine initial_conditions

  subroutine impose_user_bc(u,nb,time)
    use parameters, only : neq, nxmin, nxmax, nymin, nymax, pi, cv
    use globals, only : dx, dy, icoord, lp
    implicit none
    real,    intent(inout) :: u(neq,nxmin:nxmax,nymin:nymax)
    integer, intent(in)    :: nb
    real,    intent(in)    :: time
    integer :: lev, i, j
    real    :: x, y, vJet
    lev=lp(nb,1)
    vJet = vj0 + dvJ*sin(2.*pi*time/tauJ)
    do i=nxmin,nxmax
       do j=nymin,nymax
          x= ( float(i+icoord(nb,1)) -0.5 )* dx(lev)
          y= ( float(j+icoord(nb,2)) -0.5 )* dy(lev)
          if( (abs(x-posJx) <= LJ).and.( abs(y-posJy) <= RJ ) )  then
             u(1,i,j)= denJ
             u(2,i,j)= denJ*vJet
             u(3,i,j)= 0.
             u(4,i,j)= 0.5*denJ*vJet**2+cv*denJ*efJ*TempJ
             if (neq >= 5) u(min(5,neq),i,j)= denJ*yJ
             if (neq >= 6) u(min(6,neq),i,j)= denJ
          endif
       end do
    end do
  end subroutine impose_user_bc
end module user_mod
