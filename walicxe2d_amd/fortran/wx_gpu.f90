!=======================================================================
! wx_gpu -- iso_c_binding layer between the reference's Fortran host and
! the MI355X library (include/wx_capi.h).
!
! Build: add this file and the drop-ins of this directory to a problem
! directory ahead of ../src in VPATH (JET/Makefile:8), e.g.
!     VPATH = . : ../walicxe2d_amd/fortran : ../src
! and link with -L<repo>/walicxe2d_amd -lwxhip.  The drop-ins shadow
! tstep.f90, timestep.f90, markref.f90, refine.f90 and coarse.f90.
!
! The device holds the only up-to-date copy of u; call wx_gpu_download()
! before anything on the host reads u or primit (output.f90).
!=======================================================================
module wx_gpu
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: WX_LAYOUT_FORTRAN = 1

  type, bind(C) :: wx_params_t
     integer(c_int32_t) :: neq, nlevs, nbrootx, nbrooty
     integer(c_int32_t) :: solver, limiter, cooling, otherb
     integer(c_int32_t) :: bc(4)
     integer(c_int32_t) :: cylindrical, compat, nblocks, pad0
     real(c_double) :: cv, gamma, eta, cfl
     real(c_double) :: xmax, ymax
     real(c_double) :: Tempsc, Psc, tsc
     real(c_double) :: tol, tol2
     real(c_double) :: dx(16), dy(16)
     real(c_double) :: cooltab(82)
     real(c_double) :: jet(12)
  end type wx_params_t

  interface
     integer(c_int) function wx_create(p, ctx) bind(C, name="wx_create")
       import :: c_int, c_ptr, wx_params_t
       type(wx_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function wx_create
     integer(c_int) function wx_destroy(ctx) bind(C, name="wx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wx_destroy
     integer(c_int) function wx_upload(ctx, u, layout, lp, leaf, icoord, ld, nbmax, nbleafs) &
          bind(C, name="wx_upload")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u(*)
       integer(c_int32_t), value :: layout, ld, nbmax, nbleafs
       integer(c_int32_t), intent(in) :: lp(*), leaf(*), icoord(*)
     end function wx_upload
     integer(c_int) function wx_upload_tree(ctx, lp, leaf, icoord, ld, nbmax, nbleafs) &
          bind(C, name="wx_upload_tree")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: lp(*), leaf(*), icoord(*)
       integer(c_int32_t), value :: ld, nbmax, nbleafs
     end function wx_upload_tree
     integer(c_int) function wx_download(ctx, u, prim, layout, ld) bind(C, name="wx_download")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: u(*), prim(*)
       integer(c_int32_t), value :: layout, ld
     end function wx_download
     ! BIN checkpoints (src/output.f90:39-67 / src/basegrid.f90:167-251);
     ! paths are NUL-terminated (trim(name)//c_null_char)
     integer(c_int) function wx_write_bin(ctx, pointers, blocks, logical_true) bind(C, name="wx_write_bin")
       import :: c_int, c_ptr, c_char, c_int32_t
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: pointers(*), blocks(*)
       integer(c_int32_t), value :: logical_true
     end function wx_write_bin
     ! Hilbert placement plan of src/loadbalance.f90:5-105 (host only): keys
     ! and new block IDs of blocks 1..nbmax for nranks ranks
     integer(c_int) function wx_lb_plan(lp, leaf, icoord, ld, nbmax, nlevs, nblocks, nranks, hilbert, idnew) &
          bind(C, name="wx_lb_plan")
       import :: c_int, c_int32_t
       integer(c_int32_t), intent(in) :: lp(*), leaf(*), icoord(*)
       integer(c_int32_t), value :: ld, nbmax, nlevs, nblocks, nranks
       integer(c_int32_t), intent(out) :: hilbert(*), idnew(*)
     end function wx_lb_plan
     integer(c_int) function wx_set_bounds(ctx, ib, ld, n) bind(C, name="wx_set_bounds")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: ib(*)
       integer(c_int32_t), value :: ld, n
     end function wx_set_bounds
     integer(c_int) function wx_locatebounds(ctx, n) bind(C, name="wx_locatebounds")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: n
     end function wx_locatebounds
     integer(c_int) function wx_timestep(ctx, dt) bind(C, name="wx_timestep")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt
     end function wx_timestep
     integer(c_int) function wx_tstep(ctx, time, dt) bind(C, name="wx_tstep")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time, dt
     end function wx_tstep
     integer(c_int) function wx_criteria(ctx, up, dn) bind(C, name="wx_criteria")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: up(*), dn(*)
     end function wx_criteria
     integer(c_int) function wx_refine(ctx, nb, s1, s2, s3, s4) bind(C, name="wx_refine")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nb, s1, s2, s3, s4
     end function wx_refine
     integer(c_int) function wx_coarse(ctx, dad, s1, s2, s3, s4) bind(C, name="wx_coarse")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: dad, s1, s2, s3, s4
     end function wx_coarse
     integer(c_int) function wx_last_error(buf, n) bind(C, name="wx_last_error")
       import :: c_int, c_char, c_int32_t
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int32_t), value :: n
     end function wx_last_error
  end interface

  type(c_ptr), save :: wx_h = c_null_ptr
  logical, save :: wx_on = .false.
  ! .true.: the library builds the ghost list from its own copy of the tree
  ! (wx_locatebounds); .false.: the host's locatebounds list is handed over
  logical, save :: wx_device_bounds = .false.

contains

  !---------------------------------------------------------------------
  ! Stops with the reference's style of message on a non-zero status.
  subroutine wx_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: k, ierr
    if (rc == 0) return
    ierr = wx_last_error(buf, 512)
    msg = ' '
    do k = 1, 512
       if (buf(k) == c_null_char) exit
       msg(k:k) = buf(k)
    end do
    if (rc == 1) print '(a)', 'Error in hllc'
    print '(a,a,a,i3,a,a)', 'wx_gpu: ', what, ' failed (status', rc, '): ', trim(msg)
    stop
  end subroutine wx_check

  !---------------------------------------------------------------------
  ! Creates the device context from the problem's parameters.  The caller
  ! passes the run-time image of the cpp switches it was compiled with.
  subroutine wx_gpu_init(neq_, nlevs_, nblocks_, solver, limiter, cooling, bc, otherb, cyl, &
                         cv_, gamma_, eta_, cfl_, Tempsc_, Psc_, tsc_, tol, tol2, dx_, dy_, &
                         cooltab_, jet_)
    integer, intent(in) :: neq_, nlevs_, nblocks_, solver, limiter, cooling, bc(4), otherb, cyl
    real(kind=8), intent(in) :: cv_, gamma_, eta_, cfl_, Tempsc_, Psc_, tsc_, tol, tol2
    real(kind=8), intent(in) :: dx_(nlevs_), dy_(nlevs_), cooltab_(82), jet_(12)
    type(wx_params_t) :: p
    p%neq = neq_; p%nlevs = nlevs_; p%nbrootx = 1; p%nbrooty = 1
    p%solver = solver; p%limiter = limiter; p%cooling = cooling; p%otherb = otherb
    p%bc = bc; p%cylindrical = cyl; p%compat = 3; p%nblocks = nblocks_; p%pad0 = 0
    p%cv = cv_; p%gamma = gamma_; p%eta = eta_; p%cfl = cfl_
    p%xmax = 0.; p%ymax = 0.
    p%Tempsc = Tempsc_; p%Psc = Psc_; p%tsc = tsc_
    p%tol = tol; p%tol2 = tol2
    p%dx = 0.; p%dy = 0.
    p%dx(1:nlevs_) = dx_; p%dy(1:nlevs_) = dy_
    p%cooltab = cooltab_
    p%jet = jet_
    call wx_check(wx_create(p, wx_h), 'wx_create')
    wx_on = .true.
  end subroutine wx_gpu_init

  subroutine wx_gpu_finalize()
    integer(c_int) :: rc
    if (wx_on) rc = wx_destroy(wx_h)
    wx_on = .false.
  end subroutine wx_gpu_finalize

end module wx_gpu
