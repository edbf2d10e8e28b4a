!=======================================================================
! wx_gpu_host -- glue between the reference's module globals
! (src/parameters.f90, src/globals.f90) and the wx_gpu C-ABI layer.
! Compiled with the same cpp switches as the reference (JET/Makefile:118-209),
! which it turns into the library's run-time parameters.
!=======================================================================
module wx_gpu_host
  use iso_c_binding
  use parameters
  use globals
  use wx_gpu
  implicit none
contains

  ! Creates the device context and uploads the host state on first use.
  subroutine wx_gpu_ensure()
#ifdef COOLINGDMC
    use dmc_module, only : cooltab
#endif
#ifdef WX_JET
    use jet, only : RJ, LJ, posJ, denJ, TempJ, vJ0, dvJ, tau
#endif
    integer :: solver, limiter, cooling, bc(4), otherb, cyl
    real(kind=8) :: ct(82), jt(12)
    if (wx_on) return
    solver = 0
#ifdef HLL
    solver = 1
#endif
#ifdef HLL_HLLC
    solver = 2
#endif
    limiter = LIMITER
    cooling = 0
#ifdef COOLINGH
    cooling = 1
#endif
#ifdef COOLINGDMC
    cooling = 2
#endif
    bc = 0
#ifdef REFXL
    bc(1) = 1
#endif
#ifdef REFXR
    bc(2) = 1
#endif
#ifdef REFYB
    bc(3) = 1
#endif
#ifdef REFYT
    bc(4) = 1
#endif
    otherb = 0
    cyl = 0
#ifdef CYLINDRICAL
    cyl = 1
#endif
    ct = 0.d0
#ifdef COOLINGDMC
    ct = reshape(cooltab, (/82/))
#endif
    jt = 0.d0
#ifdef OTHERB
#ifdef WX_JET
    ! top-hat jet of JET/jet.f90 (neutral fraction 0.9999, energy factor 1.0001)
    otherb = 1
    jt = (/ dble(RJ), dble(LJ), dble(posJ(1)), dble(posJ(2)), dble(denJ), dble(TempJ), &
            dble(vJ0), dble(dvJ), dble(tau), 0.9999d0, 1.0001d0, 0.d0 /)
#else
    print '(a)', 'wx_gpu: OTHERB needs the jet source (compile with -DWX_JET); other user BCs run on the host only'
    stop
#endif
#endif
    call wx_gpu_init(neq, nlevs, nblocks, solver, limiter, cooling, bc, otherb, cyl, &
                     dble(cv), dble(gamma), dble(eta), dble(cfl), dble(Tempsc), dble(Psc), dble(tsc), &
                     0.05d0, 1.0d0, dble(dx), dble(dy), ct, jt)
    call wx_gpu_upload()
  end subroutine wx_gpu_ensure

  ! Host -> device: the whole state (u in the reference's layout) + tree.
  subroutine wx_gpu_upload()
    integer(c_int32_t), allocatable :: leafi(:)
    allocate(leafi(nblocks))
    leafi = 0
    where (leaf(1:nblocks)) leafi = 1
    nbleafs = count(leaf(nbmin:nbmax))
    call wx_check(wx_upload(wx_h, u, WX_LAYOUT_FORTRAN, lp, leafi, icoord, int(nblocks, c_int32_t), &
                            int(nbmax, c_int32_t), int(nbleafs, c_int32_t)), 'wx_upload')
    deallocate(leafi)
  end subroutine wx_gpu_upload

  ! Host tree -> device (kept tables when unchanged).
  subroutine wx_gpu_tree()
    integer(c_int32_t), allocatable :: leafi(:)
    allocate(leafi(nblocks))
    leafi = 0
    where (leaf(1:nblocks)) leafi = 1
    nbleafs = count(leaf(nbmin:nbmax))
    call wx_check(wx_upload_tree(wx_h, lp, leafi, icoord, int(nblocks, c_int32_t), &
                                 int(nbmax, c_int32_t), int(nbleafs, c_int32_t)), 'wx_upload_tree')
    deallocate(leafi)
  end subroutine wx_gpu_tree

  ! Device -> host: u and primit in the reference's layout (before output).
  subroutine wx_gpu_download()
    if (.not. wx_on) return
    call wx_check(wx_download(wx_h, u, primit, WX_LAYOUT_FORTRAN, int(nblocks, c_int32_t)), 'wx_download')
  end subroutine wx_gpu_download

end module wx_gpu_host
