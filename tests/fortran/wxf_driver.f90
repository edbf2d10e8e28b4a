!=======================================================================
! Runs N steps of the reference's time loop (src/main.f90:87-95) through
! the drop-in timestep/tstep, from a state file written by the test, and
! writes u back in the reference's layout.  TEST SCAFFOLDING.
!=======================================================================
program wxf_driver
  use parameters
  use globals
  use wx_gpu
  use wx_gpu_host
  implicit none
  integer :: nsteps, s, k, hdr(2), il
  integer, allocatable :: leafi(:)
  real :: time, dt
  character(len=32) :: arg
  call get_command_argument(1, arg)
  read(arg, *) nsteps
  rank = 0; np = 1
  nbmin = 1; nbmaxtot = nblocks
  allocate(u(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
  allocate(up(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
  allocate(primit(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
  allocate(lp(nblocks,8), leaf(nblocks), icoord(nbmin:nbmaxtot,2), leafi(nblocks))
  allocate(dx(nlevs), dy(nlevs), innerbounds(4*nblocks,3))
  do k = 1, nlevs
     dx(k) = xmax/nx/(2.**(k-1))/float(nbroot)
     dy(k) = ymax/ny/(2.**(k-1))/float(nbroot)
  end do
  open(newunit=il, file='state_in.bin', access='stream', form='unformatted', status='old')
  read(il) hdr
  read(il) u
  read(il) lp
  read(il) leafi
  read(il) icoord
  close(il)
  nbmax = hdr(1); nbleafs = hdr(2)
  leaf = leafi /= 0
  wx_device_bounds = .true.
  time = 0.
  do s = 1, nsteps
     call timestep(dt)
     time = time + dt
     call tstep(time, dt)
  end do
  call wx_gpu_download()
  open(newunit=il, file='state_out.bin', access='stream', form='unformatted', status='replace')
  write(il) time, dt
  write(il) u
  close(il)
  call wx_gpu_finalize()
end program wxf_driver
