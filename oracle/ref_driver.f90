! ref_driver.f90 -- prints the reference's 12-point quadrature grid
! (gesqua, qmc/vmc/rotqua.f90) and the grid rotated by rotqua after setrn
! with the four seed limbs given on the command line.  Test infrastructure,
! built only into oracle/_ref next to the reference's own sources.
program ref_driver
  use dim_mod, only: ndim
  use qua_mod
  implicit none
  integer :: seed(4), i, nq
  character(len=32) :: arg
  real*8 :: w(12)
  do i = 1, 4
    call get_command_argument(i, arg)
    read(arg, *) seed(i)
  enddo
  ndim = 3
  nq = 12
  allocate(xq0(12), yq0(12), zq0(12), xq(12), yq(12), zq(12), wq(12))
  call gesqua(nq, xq0, yq0, zq0, w)
  nquad = nq
  do i = 1, nq
    write(*, '(a,4es26.17e3)') 'Q ', xq0(i), yq0(i), zq0(i), w(i)
  enddo
  call setrn(seed)
  call rotqua
  do i = 1, nq
    write(*, '(a,3es26.17e3)') 'R ', xq(i), yq(i), zq(i)
  enddo
end program
