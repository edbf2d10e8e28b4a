!=======================================================================
! Drop-in for src/tstep.f90:4-126 -- one full step on the GPU.
! The host keeps locatebounds (src/locatebounds.f90) and hands its list to
! the device; stage A, boundaryII, stage B + viscosity, boundaryI and the
! cooling run as three HIP kernels over every leaf block.
!=======================================================================
subroutine tstep(time,dt)
  use iso_c_binding
  use parameters
  use globals
  use wx_gpu
  use wx_gpu_host
  implicit none
  real, intent(in) :: time, dt
  integer(c_int32_t) :: n
  call wx_gpu_ensure()
  call wx_gpu_tree()
  if (wx_device_bounds) then
     call wx_check(wx_locatebounds(wx_h, n), 'wx_locatebounds')
  else
     call locatebounds(time)
     call wx_check(wx_set_bounds(wx_h, innerbounds, int(size(innerbounds,1), c_int32_t), &
                                 int(ninternal, c_int32_t)), 'wx_set_bounds')
  end if
  call wx_check(wx_tstep(wx_h, real(time, c_double), real(dt, c_double)), 'wx_tstep')
end subroutine tstep
