! ref_shim.f90 -- bind(C) entry points over the reference's own routines
! (qmc/vmc/rannyu.f90, qmc/vmc/matinv.f90, qmc/vmc/rotqua.f90,
! qmc/dmc/gauss.f90) so that the oracle can be checked against them from
! Python.  Test infrastructure; compiled only into oracle/_ref.
subroutine ref_setrn(iseed) bind(C, name="ref_setrn")
  use iso_c_binding
  integer(c_int32_t), intent(in) :: iseed(4)
  integer :: is(4)
  is = iseed
  call setrn(is)
end subroutine

subroutine ref_savern(iseed) bind(C, name="ref_savern")
  use iso_c_binding
  integer(c_int32_t), intent(out) :: iseed(4)
  integer :: is(4)
  call savern(is)
  iseed = is
end subroutine

function ref_rannyu() result(r) bind(C, name="ref_rannyu")
  use iso_c_binding
  real(c_double) :: r
  real*8, external :: rannyu
  r = rannyu(0)
end function

function ref_gauss() result(r) bind(C, name="ref_gauss")
  use iso_c_binding
  use const_mod, only: pi
  real(c_double) :: r
  real*8, external :: gauss
  pi = 4.d0*atan(1.d0)
  r = gauss()
end function

function ref_matinv(a, n) result(det) bind(C, name="ref_matinv")
  use iso_c_binding
  integer(c_int32_t), value :: n
  real(c_double), intent(inout) :: a(n*n)
  real(c_double) :: det
  real*8 :: d
  integer :: nn
  nn = n
  call matinv(a, nn, d)
  det = d
end function

subroutine ref_gesqua(nq, x, y, z, w) bind(C, name="ref_gesqua")
  use iso_c_binding
  use dim_mod, only: ndim
  integer(c_int32_t), intent(inout) :: nq
  real(c_double), intent(out) :: x(*), y(*), z(*), w(*)
  integer :: n
  ndim = 3
  n = nq
  call gesqua(n, x, y, z, w)
  nq = n
end subroutine

subroutine ref_rotqua(nq, x0, y0, z0, x, y, z) bind(C, name="ref_rotqua")
  use iso_c_binding
  use qua_mod
  integer(c_int32_t), value :: nq
  real(c_double), intent(in) :: x0(nq), y0(nq), z0(nq)
  real(c_double), intent(out) :: x(nq), y(nq), z(nq)
  if (allocated(xq0)) deallocate(xq0, yq0, zq0, xq, yq, zq)
  allocate(xq0(nq), yq0(nq), zq0(nq), xq(nq), yq(nq), zq(nq))
  nquad = nq
  xq0 = x0; yq0 = y0; zq0 = z0
  call rotqua
  x = xq; y = yq; z = zq
end subroutine
