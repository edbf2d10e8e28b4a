!=======================================================================
! C-callable glue around two routines of the reference's load balancer,
! hilbert2d and indexx (src/loadbalance.f90:251-400).  They are internal
! procedures of the MPI-only subroutine loadbalance, so the recipe in
! Makefile extracts their text unmodified from the source file into
! oracle/_ref/ and compiles it there.  TEST INFRASTRUCTURE ONLY.
!=======================================================================
subroutine wxr_lb_hilbert2d(x, y, order, bits, n) bind(C, name="wxr_lb_hilbert2d")
  use iso_c_binding
  implicit none
  integer(c_int), value :: bits, n
  integer(c_int), intent(in) :: x(n), y(n)
  integer(c_int), intent(out) :: order(n)
  call hilbert2d(x, y, order, bits, n)
end subroutine wxr_lb_hilbert2d

subroutine wxr_lb_indexx(n, arr, indx) bind(C, name="wxr_lb_indexx")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  integer(c_int), intent(in) :: arr(n)
  integer(c_int), intent(out) :: indx(n)
  call indexx(n, arr, indx)
end subroutine wxr_lb_indexx
