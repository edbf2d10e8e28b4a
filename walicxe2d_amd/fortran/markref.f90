!=======================================================================
! Drop-in for src/markref.f90:4-85 -- criteria (src/criteria.f90:4-53) of
! every leaf evaluated on the device in one call.
!=======================================================================
subroutine markref(irefup,irefdown)
  use iso_c_binding
  use parameters
  use globals
  use wx_gpu
  use wx_gpu_host
  implicit none
  logical, intent(out), dimension(nblocks*np) :: irefup, irefdown
  integer(c_int32_t), allocatable :: fup(:), fdn(:)
  call wx_gpu_ensure()
  call wx_gpu_tree()
  allocate(fup(nblocks), fdn(nblocks))
  call wx_check(wx_criteria(wx_h, fup, fdn), 'wx_criteria')
  irefup(:) = .false.
  irefdown(:) = .false.
  irefup(1:nblocks) = fup(1:nblocks) /= 0
  irefdown(1:nblocks) = fdn(1:nblocks) /= 0
  deallocate(fup, fdn)
end subroutine markref
