!=======================================================================
! The reference's own host side driving the GPU path: its admesh
! (src/admesh.f90, with its critneighup / critneighdown / updatelpup /
! updatelpdown) and locatebounds (src/locatebounds.f90) compiled from its
! sources, with the drop-ins tstep / timestep / markref / refine / coarse of
! walicxe2d_amd/fortran in place of the CPU routines (the reference's own
! VPATH shadowing), run through N iterations of its time loop
! (src/main.f90:87-98: timestep, time += dt, tstep, admesh).  tstep hands
! the host's locatebounds list to the device (wx_set_bounds).  Reads
! state_in.bin (nbmax, nbleafs, dx, dy, the DMC table, u in the reference's
! layout, lp, leaf, icoord) and writes state_out.bin (time, dt, u, lp, leaf,
! icoord, nbmax, nbleafs).  TEST SCAFFOLDING.
!=======================================================================
program wxf_amr_driver
  use parameters
  use globals
  use dmc_module
  use wx_gpu
  use wx_gpu_host
  implicit none
  integer :: nsteps, s, hdr(2), il, err
  integer, allocatable :: leafi(:)
  real :: time, dt
  character(len=32) :: arg
  call get_command_argument(1, arg)
  read(arg, *) nsteps
  call mpi_init(err)
  rank = 0; np = 1
  nbmin = 1; nbmaxtot = nblocks
  allocate(u(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
  allocate(up(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
  allocate(primit(nbmin:nbmaxtot,neq,nxmin:nxmax,nymin:nymax))
  allocate(lp(nblocks,8), leaf(nblocks), icoord(nbmin:nbmaxtot,2), leafi(nblocks))
  allocate(dx(nlevs), dy(nlevs), innerbounds(4*nblocks,3))
  allocate(extbounds(1,5))
  nexternal = 0
  open(newunit=il, file='state_in.bin', access='stream', form='unformatted', status='old')
  read(il) hdr
  read(il) dx, dy
  read(il) cooltab
  read(il) u
  read(il) lp
  read(il) leafi
  read(il) icoord
  close(il)
  nbmax = hdr(1); nbleafs = hdr(2)
  leaf = leafi /= 0
  wx_device_bounds = .false.
  time = 0.
  do s = 1, nsteps
     call timestep(dt)
     time = time + dt
     call tstep(time, dt)
     call admesh
  end do
  call wx_gpu_download()
  leafi = 0
  where (leaf) leafi = 1
  open(newunit=il, file='state_out.bin', access='stream', form='unformatted', status='replace')
  write(il) time, dt
  write(il) u
  write(il) lp
  write(il) leafi
  write(il) icoord
  write(il) nbmax, nbleafs
  close(il)
  call wx_gpu_finalize()
  call mpi_finalize(err)
end program wxf_amr_driver
