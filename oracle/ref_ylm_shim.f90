! bind(C) entry points to the reference's recursive real spherical
! harmonics (module real_spherical_harmonics and the routines after it in
! qmc/vmc/basis_fns.f90), compiled from the reference's own sources by
! ref_closure.py.  TEST INFRASTRUCTURE: used only to check the oracle.
module ref_ylm_shim
  use iso_c_binding
  use real_spherical_harmonics
  implicit none
contains
  subroutine ref_ylm_setup() bind(C, name="ref_ylm_setup")
    call setup_spherical_harmonics
    call setup_coefficients_ylm
  end subroutine

  function ref_ylm_lmax() bind(C, name="ref_ylm_lmax") result(l)
    integer(c_int) :: l
    l = lmax
  end function

  ! Y_lm r^l = coef_ylm(l,|m|) P_lm(x,y,z) and its gradient, as basis_fns
  ! forms them (qmc/vmc/basis_fns.f90:404-431), for l <= lmax, m = -l..l:
  ! p(m + lmax + 1, l + 1), dp(k, m + lmax + 1, l + 1)
  subroutine ref_ylm_eval(x, y, z, p, dp) bind(C, name="ref_ylm_eval")
    real(c_double), value :: x, y, z
    real(c_double), intent(out) :: p(-lmax:lmax, 0:lmax), dp(3, -lmax:lmax, 0:lmax)
    integer :: k, l, m
    r_inv_power_minus(:) = 1.d0
    x_power_of(:) = 0.d0; y_power_of(:) = 0.d0; z_power_of(:) = 0.d0
    x_power_of(0) = 1.d0; y_power_of(0) = 1.d0; z_power_of(0) = 1.d0
    do k = 1, lmax
      x_power_of(k) = x_power_of(k - 1) * x
      y_power_of(k) = y_power_of(k - 1) * y
      z_power_of(k) = z_power_of(k - 1) * z
    end do
    call calculate_spherical_harmonics
    p = 0.d0
    dp = 0.d0
    do l = 0, lmax
      do m = -l, l
        p(m, l) = large_p(l, m) * coef_ylm(l, abs(m))
        dp(:, m, l) = d_large_p(:, l, m) * coef_ylm(l, abs(m))
      end do
    end do
  end subroutine
end module ref_ylm_shim
