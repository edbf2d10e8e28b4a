!=======================================================================
! Drop-in for src/coarse.f90:5-368 -- 2x2 restriction of the four sons
! (interiors and ghosts) into the father on the device.
!=======================================================================
subroutine coarse(dad,son1,son2,son3,son4)
  use iso_c_binding
  use parameters
  use globals
  use wx_gpu
  use wx_gpu_host
  implicit none
  integer, intent(in) :: dad,son1,son2,son3,son4
  call wx_gpu_ensure()
  call wx_check(wx_coarse(wx_h, int(dad,c_int32_t), int(son1,c_int32_t), int(son2,c_int32_t), &
                          int(son3,c_int32_t), int(son4,c_int32_t)), 'wx_coarse')
end subroutine coarse
