!=======================================================================
! Drop-in for src/refine.f90:4-263 -- injection of the padded parent block
! into its four sons on the device; icoord stays on the host
! (src/refine.f90:227-261).
!=======================================================================
subroutine refine(nb,son1,son2,son3,son4)
  use iso_c_binding
  use parameters
  use globals
  use wx_gpu
  use wx_gpu_host
  implicit none
  integer, intent(in) :: nb,son1,son2,son3,son4
  integer :: nx0, ny0
  call wx_gpu_ensure()
  call wx_check(wx_refine(wx_h, int(nb,c_int32_t), int(son1,c_int32_t), int(son2,c_int32_t), &
                          int(son3,c_int32_t), int(son4,c_int32_t)), 'wx_refine')
  nx0 = icoord(nb,1)*2
  ny0 = icoord(nb,2)*2
  icoord(son1,1) = nx0;     icoord(son1,2) = ny0
  icoord(son2,1) = nx0+nx;  icoord(son2,2) = ny0
  icoord(son3,1) = nx0;     icoord(son3,2) = ny0+ny
  icoord(son4,1) = nx0+nx;  icoord(son4,2) = ny0+ny
end subroutine refine
