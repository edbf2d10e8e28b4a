!=======================================================================
! C-callable glue around the reference's own hot-path routines, used to
! write golden vectors and to check the C oracle.  TEST INFRASTRUCTURE
! ONLY.  C-side arrays use the layout [id][var][j][i] (i fastest, ID 0
! unused), i.e. Fortran cu(-1:18,-1:18,neq,0:cap); the reference's own
! arrays keep their block-fastest layout u(nb,ieq,i,j).
!=======================================================================
module wxref_glue
  use iso_c_binding
  use parameters
  use globals
  implicit none
  logical, save :: mpi_up = .false.
contains

  subroutine wxr_init(cap) bind(C, name="wxr_init")
    integer(c_int), value :: cap
    integer :: err
    logical :: flag
#ifdef MPIP
    if (.not. mpi_up) then
       call mpi_initialized(flag, err)
       if (.not. flag) call mpi_init(err)
       mpi_up = .true.
    end if
#endif
    rank = 0
    np = 1
    nbmin = 1
    nbmaxtot = nblocks
    nbmax = 0
    nbleafs = 0
    if (allocated(u)) deallocate(u, up, primit, lp, leaf, icoord, dx, dy, innerbounds)
    allocate(u(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
    allocate(up(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
    allocate(primit(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
    allocate(lp(nblocks*np,8), leaf(nblocks*np), icoord(nbmin:nbmaxtot,2))
    allocate(dx(nlevs), dy(nlevs))
    allocate(innerbounds(4*nblocks,3))
#ifdef MPIP
    if (allocated(extbounds)) deallocate(extbounds)
    allocate(extbounds(1,5))
    nexternal = 0
#endif
    u = 0.; up = 0.; primit = 0.
    lp = 0; leaf = .false.; icoord = 0
    innerbounds = 0
    ninternal = 0
  end subroutine wxr_init

  subroutine wxr_consts(out) bind(C, name="wxr_consts")
    real(c_double), intent(out) :: out(12)
    out(1) = cv; out(2) = gamma; out(3) = Tempsc; out(4) = Psc; out(5) = tsc
    out(6) = cfl; out(7) = eta; out(8) = real(neq); out(9) = real(nlevs)
    out(10) = real(nblocks); out(11) = rsc; out(12) = vsc2
  end subroutine wxr_consts

  subroutine wxr_set_dxdy(cdx, cdy) bind(C, name="wxr_set_dxdy")
    real(c_double), intent(in) :: cdx(nlevs), cdy(nlevs)
    dx(:) = cdx(:)
    dy(:) = cdy(:)
  end subroutine wxr_set_dxdy

  subroutine wxr_put(cap, cu, cup, cprim, clp, cleaf, cic, cnbmax, cnbleafs) bind(C, name="wxr_put")
    integer(c_int), value :: cap, cnbmax, cnbleafs
    real(c_double), intent(in) :: cu(-1:18,-1:18,neq,0:cap), cup(-1:18,-1:18,neq,0:cap)
    real(c_double), intent(in) :: cprim(-1:18,-1:18,neq,0:cap)
    integer(c_int), intent(in) :: clp(8,0:cap), cleaf(0:cap), cic(2,0:cap)
    integer :: nb, k, i, j, top
    top = min(cap, nblocks)
    do nb = 1, top
       do k = 1, neq
          do j = -1, 18
             do i = -1, 18
                u(nb,k,i,j) = cu(i,j,k,nb)
                up(nb,k,i,j) = cup(i,j,k,nb)
                primit(nb,k,i,j) = cprim(i,j,k,nb)
             end do
          end do
       end do
       lp(nb,:) = clp(:,nb)
       leaf(nb) = (cleaf(nb) /= 0)
       icoord(nb,:) = cic(:,nb)
    end do
    nbmax = cnbmax
    nbleafs = cnbleafs
  end subroutine wxr_put

  subroutine wxr_get(cap, cu, cup, cprim, clp, cleaf, cic, cnbmax, cnbleafs) bind(C, name="wxr_get")
    integer(c_int), value :: cap
    real(c_double), intent(out) :: cu(-1:18,-1:18,neq,0:cap), cup(-1:18,-1:18,neq,0:cap)
    real(c_double), intent(out) :: cprim(-1:18,-1:18,neq,0:cap)
    integer(c_int), intent(out) :: clp(8,0:cap), cleaf(0:cap), cic(2,0:cap)
    integer(c_int), intent(out) :: cnbmax, cnbleafs
    integer :: nb, k, i, j, top
    top = min(cap, nblocks)
    do nb = 1, top
       do k = 1, neq
          do j = -1, 18
             do i = -1, 18
                cu(i,j,k,nb) = u(nb,k,i,j)
                cup(i,j,k,nb) = up(nb,k,i,j)
                cprim(i,j,k,nb) = primit(nb,k,i,j)
             end do
          end do
       end do
       clp(:,nb) = lp(nb,:)
       cleaf(nb) = merge(1, 0, leaf(nb))
       cic(:,nb) = icoord(nb,:)
    end do
    cnbmax = nbmax
    cnbleafs = nbleafs
  end subroutine wxr_get

  ! install a given list as innerbounds (the reference's locatebounds
  ! reallocates innerbounds to nbleafs*4 rows; grow it back when needed)
  subroutine wxr_set_bounds(cb, n) bind(C, name="wxr_set_bounds")
    integer(c_int), value :: n
    integer(c_int), intent(in) :: cb(3,n)
    integer :: k
    if (size(innerbounds, 1) < max(n, 1)) then
       deallocate(innerbounds)
       allocate(innerbounds(max(n, 4*nblocks),3))
    end if
    do k = 1, n
       innerbounds(k,:) = cb(:,k)
    end do
    ninternal = n
  end subroutine wxr_set_bounds

  ! the reference's own locatebounds; returns ninternal, copies the list
  function wxr_locatebounds(time, cb, nmax) bind(C, name="wxr_locatebounds") result(n)
    real(c_double), value :: time
    integer(c_int), value :: nmax
    integer(c_int), intent(out) :: cb(3,nmax)
    integer(c_int) :: n
    real :: t
    integer :: k
    t = time
    call locatebounds(t)
    n = ninternal
    do k = 1, min(n, nmax)
       cb(:,k) = innerbounds(k,:)
    end do
  end function wxr_locatebounds

  ! the reference's own output (src/output.f90, OUTBIN branch): BIN/pointers
  ! and BIN/blocks files of rank 0 under the working directory
  subroutine wxr_output(itprint) bind(C, name="wxr_output")
    integer(c_int), value :: itprint
    integer :: it
    it = itprint
    call output(it)
  end subroutine wxr_output

  ! leaf blocks of the reference's mesh (its nbleafs global)
  function wxr_nbleafs() bind(C, name="wxr_nbleafs") result(n)
    integer(c_int) :: n
    n = nbleafs
  end function wxr_nbleafs

  ! the reference's own admesh (markref, critneighup, refine by level,
  ! critneighdown, coarsen by level)
  subroutine wxr_admesh() bind(C, name="wxr_admesh")
    call admesh
  end subroutine wxr_admesh

  subroutine wxr_set_jet(cj) bind(C, name="wxr_set_jet")
    use user_mod
    real(c_double), intent(in) :: cj(12)
    RJ = cj(1); LJ = cj(2); posJx = cj(3); posJy = cj(4); denJ = cj(5); TempJ = cj(6)
    vJ0 = cj(7); dvJ = cj(8); tauJ = cj(9); yJ = cj(10); efJ = cj(11)
  end subroutine wxr_set_jet

  subroutine wxr_set_cooltab(ct) bind(C, name="wxr_set_cooltab")
#ifdef COOLINGDMC
    use dmc_module
#endif
    real(c_double), intent(in) :: ct(2,41)
#ifdef COOLINGDMC
    cooltab(:,:) = ct(:,:)
#endif
  end subroutine wxr_set_cooltab

  function wxr_cooldmc(T) bind(C, name="wxr_cooldmc") result(r)
#ifdef COOLINGDMC
    use dmc_module
#endif
    real(c_double), value :: T
    real(c_double) :: r
    real :: tt
    tt = T
#ifdef COOLINGDMC
    r = cooldmc(tt)
#else
    r = 0.
#endif
  end function wxr_cooldmc

  subroutine wxr_uprim(cprim, cuu, cT) bind(C, name="wxr_uprim")
    real(c_double), intent(out) :: cprim(neq), cT
    real(c_double), intent(in) :: cuu(neq)
    real :: pr(neq), uu(neq), T
    uu = cuu
    call uprim(pr, uu, T)
    cprim = pr
    cT = T
  end subroutine wxr_uprim

  subroutine wxr_primu(cprim, cuu) bind(C, name="wxr_primu")
    real(c_double), intent(in) :: cprim(neq)
    real(c_double), intent(out) :: cuu(neq)
    real :: pr(neq), uu(neq)
    pr = cprim
    call primu(pr, uu)
    cuu = uu
  end subroutine wxr_primu

  subroutine wxr_primf(cprim, cff) bind(C, name="wxr_primf")
    real(c_double), intent(in) :: cprim(neq)
    real(c_double), intent(out) :: cff(neq)
    real :: pr(neq), ff(neq)
    pr = cprim
    call primf(pr, ff)
    cff = ff
  end subroutine wxr_primf

  subroutine wxr_limiter(pll, pl, pr, prr) bind(C, name="wxr_limiter")
    real(c_double), intent(in) :: pll(neq), prr(neq)
    real(c_double), intent(inout) :: pl(neq), pr(neq)
    real :: a(neq), b(neq), c(neq), d(neq)
    a = pll; b = pl; c = pr; d = prr
    call limiter(a, b, c, d, neq)
    pl = b; pr = c
  end subroutine wxr_limiter

  subroutine wxr_fluxes(nb, choice, cf, cg) bind(C, name="wxr_fluxes")
    integer(c_int), value :: nb, choice
    real(c_double), intent(out) :: cf(-1:18,-1:18,neq), cg(-1:18,-1:18,neq)
    real, dimension(neq,nxmin:nxmax,nymin:nymax) :: f, g
    integer :: k
    f = 0.; g = 0.
#ifdef HLLC
    call hllcfluxes(nb, choice, f, g)
#endif
#ifdef HLL
    call hllfluxes(nb, choice, f, g)
#endif
#ifdef HLL_HLLC
    call hll_hllcfluxes(nb, choice, f, g)
#endif
    do k = 1, neq
       cf(:,:,k) = f(k,:,:)
       cg(:,:,k) = g(k,:,:)
    end do
  end subroutine wxr_fluxes

  subroutine wxr_stepblock(nb, dt, cf, cg) bind(C, name="wxr_stepblock")
    integer(c_int), value :: nb
    real(c_double), value :: dt
    real(c_double), intent(in) :: cf(-1:18,-1:18,neq), cg(-1:18,-1:18,neq)
    real, dimension(neq,nxmin:nxmax,nymin:nymax) :: f, g
    integer :: k
    real :: dtt
    do k = 1, neq
       f(k,:,:) = cf(:,:,k)
       g(k,:,:) = cg(:,:,k)
    end do
    dtt = dt
    call step(nb, dtt, f, g)
  end subroutine wxr_stepblock

  subroutine wxr_viscosity(nb) bind(C, name="wxr_viscosity")
    integer(c_int), value :: nb
    call viscosity(nb, nbmin, nbmaxtot, u, up)
  end subroutine wxr_viscosity

  subroutine wxr_boundary(time, which) bind(C, name="wxr_boundary")
    real(c_double), value :: time
    integer(c_int), value :: which
    real :: t
    t = time
    if (which == 2) then
       call boundaryII(t)
    else
       call boundaryI(t)
    end if
  end subroutine wxr_boundary

  subroutine wxr_atomic(dt, cuu) bind(C, name="wxr_atomic")
    real(c_double), value :: dt
    real(c_double), intent(inout) :: cuu(neq)
    real :: uu(neq), d
    uu = cuu
    d = dt
#ifdef COOLINGH
    call atomic(d, uu, 1.)
#endif
    cuu = uu
  end subroutine wxr_atomic

  function wxr_aloss(x1, x2, dt, den, dh0, te0) bind(C, name="wxr_aloss") result(r)
    real(c_double), value :: x1, x2, dt, den, dh0, te0
    real(c_double) :: r
#ifdef COOLINGH
    real(kind=8) :: ALOSS
    real :: d
    d = dt
    r = ALOSS(x1, x2, d, den, dh0, te0)
#else
    r = 0.
#endif
  end function wxr_aloss

  subroutine wxr_cooling(dt) bind(C, name="wxr_cooling")
    real(c_double), value :: dt
    real :: d
    d = dt
#ifdef COOLINGH
    call coolingh(d)
#endif
#ifdef COOLINGDMC
    call coolingdmc(d)
#endif
  end subroutine wxr_cooling

  subroutine wxr_tstep(time, dt) bind(C, name="wxr_tstep")
    real(c_double), value :: time, dt
    real :: t, d
    t = time
    d = dt
    call tstep(t, d)
  end subroutine wxr_tstep

  function wxr_timestep() bind(C, name="wxr_timestep") result(r)
    real(c_double) :: r
    real :: dt
    call timestep(dt)
    r = dt
  end function wxr_timestep

  subroutine wxr_criteria(nb, cup, cdn) bind(C, name="wxr_criteria")
    integer(c_int), value :: nb
    integer(c_int), intent(inout) :: cup(0:nblocks), cdn(0:nblocks)
    logical :: iu(nblocks*np), idn(nblocks*np)
    iu = (cup(1:nblocks) /= 0)
    idn = (cdn(1:nblocks) /= 0)
    call criteria(nb, iu, idn)
    cup(1:nblocks) = merge(1, 0, iu)
    cdn(1:nblocks) = merge(1, 0, idn)
  end subroutine wxr_criteria

  subroutine wxr_markref(cup, cdn) bind(C, name="wxr_markref")
    integer(c_int), intent(inout) :: cup(0:nblocks), cdn(0:nblocks)
    logical :: iu(nblocks*np), idn(nblocks*np)
    call markref(iu, idn)
    call critneighup(iu)
    cup(1:nblocks) = merge(1, 0, iu)
    cdn(1:nblocks) = merge(1, 0, idn)
  end subroutine wxr_markref

  subroutine wxr_critneighdown(cdn) bind(C, name="wxr_critneighdown")
    integer(c_int), intent(inout) :: cdn(0:nblocks)
    logical :: idn(nblocks*np)
    idn = (cdn(1:nblocks) /= 0)
    call critneighdown(idn)
    cdn(1:nblocks) = merge(1, 0, idn)
  end subroutine wxr_critneighdown

  subroutine wxr_refine(nb, s1, s2, s3, s4, cup, cdn) bind(C, name="wxr_refine")
    integer(c_int), value :: nb, s1, s2, s3, s4
    integer(c_int), intent(inout) :: cup(0:nblocks), cdn(0:nblocks)
    logical :: iu(nblocks*np), idn(nblocks*np)
    iu = (cup(1:nblocks) /= 0)
    idn = (cdn(1:nblocks) /= 0)
    call refine(nb, s1, s2, s3, s4)
    call updatelpup(nb, s1, s2, s3, s4, iu, idn)
    cup(1:nblocks) = merge(1, 0, iu)
    cdn(1:nblocks) = merge(1, 0, idn)
  end subroutine wxr_refine

  subroutine wxr_coarse(dad, s1, s2, s3, s4) bind(C, name="wxr_coarse")
    integer(c_int), value :: dad, s1, s2, s3, s4
    call coarse(dad, s1, s2, s3, s4)
  end subroutine wxr_coarse
end module wxref_glue

