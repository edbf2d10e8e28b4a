!=======================================================================
! Drop-in for src/timestep.f90:7-61 -- CFL (+ DMC cooling-time) timestep
! reduced on the device.
!=======================================================================
subroutine timestep(dt)
  use iso_c_binding
  use parameters
  use globals
  use wx_gpu
  use wx_gpu_host
  implicit none
  real, intent(out) :: dt
  real(c_double) :: d
  call wx_gpu_ensure()
  call wx_gpu_tree()
  call wx_check(wx_timestep(wx_h, d), 'wx_timestep')
  dt = d
end subroutine timestep
