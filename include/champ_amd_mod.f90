! champ_amd_mod.f90 -- Fortran 2003 binding (iso_c_binding) of the C ABI in
! champ_amd.h, for CHAMP's Fortran drivers.  Every interface here binds one
! C entry point by name; the derived type qmcw_system is interoperable with
! struct qmcw_system (same members in the same order).
module champ_amd_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: QMCW_OK = 0, QMCW_ERR_ARG = 1, QMCW_ERR_HIP = 2, &
                               QMCW_ERR_SINGULAR = 3, QMCW_ERR_NAN = 4, QMCW_ERR_NOMEM = 5
  integer(c_int32_t), parameter :: QMCW_FP64 = 0, QMCW_MIXED = 1, QMCW_MIXED_BF16 = 2
  integer(c_int32_t), parameter :: QMCW_BASIS_SLATER = 1, QMCW_BASIS_GAUSSIAN = 2

  type, bind(C) :: qmcw_system
    integer(c_int32_t) :: nelec, nup
    integer(c_int32_t) :: ncent, nctype
    type(c_ptr) :: cent, iwctype, znuc
    integer(c_int32_t) :: basis_type, nrad
    type(c_ptr) :: rad_n, rad_prim_off, prim_zeta, prim_coef
    type(c_ptr) :: ctype_bas_off, bas_rad, bas_l, bas_m
    integer(c_int32_t) :: nbasis
    integer(c_int32_t) :: norb
    type(c_ptr) :: coef
    integer(c_int32_t) :: ndet
    type(c_ptr) :: det_up, det_dn
    integer(c_int32_t) :: ncsf
    type(c_ptr) :: csf_coef, csf_det_off, csf_det, csf_det_coef
    integer(c_int32_t) :: isc, nspin2b, nocuspb
    integer(c_int32_t) :: norda, nordb, nordc
    real(c_double) :: scalek
    type(c_ptr) :: a4, b, c
    integer(c_int32_t) :: nloc
    type(c_ptr) :: npotd, lpotp1, ecp_off, ecp_coef, ecp_power, ecp_expo
    integer(c_int32_t) :: nquad
    integer(c_int32_t) :: nparma_tot
    type(c_ptr) :: jas_a_type, jas_a_idx
    integer(c_int32_t) :: nparmb_tot
    type(c_ptr) :: jas_b_spin, jas_b_idx
    integer(c_int32_t) :: nparmc_tot
    type(c_ptr) :: jas_c_type, jas_c_idx
    integer(c_int32_t) :: nparmcsf
    type(c_ptr) :: csf_parm
    integer(c_int32_t) :: nparmorb
    type(c_ptr) :: orb_parm_i, orb_parm_j, orb_parm_rev
    real(c_double) :: hb
    type(c_ptr) :: rad_num
    integer(c_int32_t) :: nnum
    type(c_ptr) :: num_igrid, num_nr, num_h, num_r0, num_off, num_rwf, num_d2, num_ce, num_ae
  end type qmcw_system

  ! struct qmcw_dmc_params: CHAMP's DMC controls of one generation
  type, bind(C) :: qmcw_dmc_params
    real(c_double) :: tau, etrial, eest, esigma, pop_ctrl, taueff
    real(c_double) :: c_rewt, pow_rewt, limit_rewt, adrift
    integer(c_int32_t) :: n_drift, itau_eff, itau_integ, iacc_rej, icross, ene_int, idmc
    integer(c_int32_t) :: tmoves
  end type qmcw_dmc_params

  ! struct qmcw_dmc_stats: population sums of one generation
  type, bind(C) :: qmcw_dmc_stats
    real(c_double) :: wsum, wesum, we2sum, wtausum, pacc, nacc, ntry, ncross
    integer(c_int32_t) :: nwalk, nbranch
  end type qmcw_dmc_stats

  interface
    function qmcw_sizeof_system() bind(C, name="qmcw_sizeof_system") result(n)
      import :: c_int64_t
      integer(c_int64_t) :: n
    end function
    function qmcw_nterms4(nordc) bind(C, name="qmcw_nterms4") result(n)
      import :: c_int32_t
      integer(c_int32_t), value :: nordc
      integer(c_int32_t) :: n
    end function
    function qmcw_create(sys, nwalk, precision, h) bind(C, name="qmcw_create") result(st)
      import :: qmcw_system, c_int32_t, c_ptr, c_int
      type(qmcw_system), intent(in) :: sys
      integer(c_int32_t), value :: nwalk, precision
      type(c_ptr), intent(out) :: h
      integer(c_int) :: st
    end function
    function qmcw_destroy(h) bind(C, name="qmcw_destroy") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: st
    end function
    function qmcw_set_params(h, a4, b, c, csf_coef, coef) bind(C, name="qmcw_set_params") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: h, a4, b, c, csf_coef, coef
      integer(c_int) :: st
    end function
    function qmcw_set_walkers(h, x) bind(C, name="qmcw_set_walkers") result(st)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*)
      integer(c_int) :: st
    end function
    function qmcw_get_walkers(h, x) bind(C, name="qmcw_get_walkers") result(st)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: x(*)
      integer(c_int) :: st
    end function
    function qmcw_eval_full(h, logpsid, sgn, psij, grad, lap, eloc, epot) &
        bind(C, name="qmcw_eval_full") result(st)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: logpsid(*), sgn(*), psij(*), grad(*), lap(*), eloc(*), epot(*)
      integer(c_int) :: st
    end function
    function qmcw_ratio_1e(h, ie, xnew, ratio, grad, lap) bind(C, name="qmcw_ratio_1e") result(st)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: ie
      real(c_double), intent(in) :: xnew(*)
      real(c_double), intent(out) :: ratio(*), grad(*), lap(*)
      integer(c_int) :: st
    end function
    function qmcw_accept_1e(h, ie, accept) bind(C, name="qmcw_accept_1e") result(st)
      import :: c_ptr, c_int, c_int32_t, c_int8_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: ie
      integer(c_int8_t), intent(in) :: accept(*)
      integer(c_int) :: st
    end function
    function qmcw_grad_1e(h, ie, grad, lap) bind(C, name="qmcw_grad_1e") result(st)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: ie
      real(c_double), intent(out) :: grad(*), lap(*)
      integer(c_int) :: st
    end function
    function qmcw_nparm(h) bind(C, name="qmcw_nparm") result(n)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: n
    end function
    function qmcw_derivs(h, dpsi, deloc) bind(C, name="qmcw_derivs") result(st)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: dpsi(*), deloc(*)
      integer(c_int) :: st
    end function
    function qmcw_sample_derivs(h, enable) bind(C, name="qmcw_sample_derivs") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: enable
      integer(c_int) :: st
    end function
    function qmcw_get_derivs(h, dpsi, deloc) bind(C, name="qmcw_get_derivs") result(st)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: dpsi(*), deloc(*)
      integer(c_int) :: st
    end function
    ! sums of the linear method's derivative products (qmcw_accum_get layout:
    ! 3 + 3 nparm + 3 nparm**2 doubles)
    function qmcw_accum_reset(h) bind(C, name="qmcw_accum_reset") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: st
    end function
    function qmcw_accum_add(h) bind(C, name="qmcw_accum_add") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: st
    end function
    function qmcw_accum_get(h, acc) bind(C, name="qmcw_accum_get") result(st)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: acc(*)
      integer(c_int) :: st
    end function
    function qmcw_set_rng(h, seed) bind(C, name="qmcw_set_rng") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(in) :: seed(*)
      integer(c_int) :: st
    end function
    function qmcw_get_rng(h, seed) bind(C, name="qmcw_get_rng") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: seed(*)
      integer(c_int) :: st
    end function
    function qmcw_vmc_run(h, nstep, delta, fbias, eloc_trace, naccept) &
        bind(C, name="qmcw_vmc_run") result(st)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nstep
      real(c_double), value :: delta, fbias
      type(c_ptr), value :: eloc_trace, naccept
      integer(c_int) :: st
    end function
    function qmcw_vmc_run_device(h, nstep, delta, fbias) bind(C, name="qmcw_vmc_run_device") result(st)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nstep
      real(c_double), value :: delta, fbias
      integer(c_int) :: st
    end function
    function qmcw_batch_create(hs, n, b) bind(C, name="qmcw_batch_create") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), intent(in) :: hs(*)
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: b
      integer(c_int) :: st
    end function
    function qmcw_batch_vmc_run(b, nstep, delta, fbias) bind(C, name="qmcw_batch_vmc_run") result(st)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: b
      integer(c_int32_t), value :: nstep
      real(c_double), value :: delta, fbias
      integer(c_int) :: st
    end function
    function qmcw_batch_destroy(b) bind(C, name="qmcw_batch_destroy") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: b
      integer(c_int) :: st
    end function
    function qmcw_batch_last_error(b) bind(C, name="qmcw_batch_last_error") result(msg)
      import :: c_ptr
      type(c_ptr), value :: b
      type(c_ptr) :: msg
    end function
    function qmcw_get_eloc(h, eloc) bind(C, name="qmcw_get_eloc") result(st)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: eloc(*)
      integer(c_int) :: st
    end function
    function qmcw_sync(h) bind(C, name="qmcw_sync") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: st
    end function
    function qmcw_last_error(h) bind(C, name="qmcw_last_error") result(msg)
      import :: c_ptr
      type(c_ptr), value :: h
      type(c_ptr) :: msg
    end function
    function qmcw_autotune(h, chosen) bind(C, name="qmcw_autotune") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: chosen
      integer(c_int) :: st
    end function
    ! DMC (qmc/rundmc_mod.f90): weight may be c_null_ptr (all 1)
    function qmcw_dmc_begin(h, nwalk, weight) bind(C, name="qmcw_dmc_begin") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nwalk
      type(c_ptr), value :: weight
      integer(c_int) :: st
    end function
    function qmcw_dmc_step(h, p, stats) bind(C, name="qmcw_dmc_step") result(st)
      import :: c_ptr, c_int, qmcw_dmc_params, qmcw_dmc_stats
      type(c_ptr), value :: h
      type(qmcw_dmc_params), intent(in) :: p
      type(qmcw_dmc_stats), intent(out) :: stats
      integer(c_int) :: st
    end function
    function qmcw_dmc_get(h, nwalk, weight) bind(C, name="qmcw_dmc_get") result(st)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: nwalk
      real(c_double), intent(out) :: weight(*)
      integer(c_int) :: st
    end function
    function qmcw_set_branch_rng(h, seed) bind(C, name="qmcw_set_branch_rng") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(in) :: seed(*)
      integer(c_int) :: st
    end function
    function qmcw_get_branch_rng(h, seed) bind(C, name="qmcw_get_branch_rng") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: seed(*)
      integer(c_int) :: st
    end function
    ! DMC branching of a host population (split_join_mod.f90:3-219), on the
    ! host or on the device in one workgroup
    function qmcw_split_join(n, mwalk, weight, origin, rng, brng, nwalk_out, nbranch) &
        bind(C, name="qmcw_split_join") result(st)
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: n, mwalk
      real(c_double), intent(inout) :: weight(*)
      integer(c_int32_t), intent(inout) :: origin(*), rng(*), brng(*)
      integer(c_int32_t), intent(out) :: nwalk_out, nbranch
      integer(c_int) :: st
    end function
    function qmcw_split_join_gpu(n, mwalk, weight, origin, rng, brng, nwalk_out, nbranch) &
        bind(C, name="qmcw_split_join_gpu") result(st)
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: n, mwalk
      real(c_double), intent(inout) :: weight(*)
      integer(c_int32_t), intent(inout) :: origin(*), rng(*), brng(*)
      integer(c_int32_t), intent(out) :: nwalk_out, nbranch
      integer(c_int) :: st
    end function
    ! VMC mover (imetro 1: metrop_mov1, 6: metrop_slat_mov1) and its steps
    function qmcw_set_vmc_mover(h, imetro, deltar, deltat) bind(C, name="qmcw_set_vmc_mover") result(st)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: imetro
      real(c_double), value :: deltar, deltat
      integer(c_int) :: st
    end function
    ! full re-inversion of the Slater matrices every `every` moves (0: never)
    function qmcw_set_reinversion(h, every) bind(C, name="qmcw_set_reinversion") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: every
      integer(c_int) :: st
    end function
    function qmcw_reinvert(h) bind(C, name="qmcw_reinvert") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: st
    end function
    ! derivative-product sums into a caller-held device buffer, no allocation
    function qmcw_accum_add_static(h, acc) bind(C, name="qmcw_accum_add_static") result(st)
      import :: c_ptr, c_int
      type(c_ptr), value :: h, acc
      integer(c_int) :: st
    end function
    ! lane-group widths and per-kernel timing
    function qmcw_set_group_width(h, g) bind(C, name="qmcw_set_group_width") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: g
      integer(c_int) :: st
    end function
    function qmcw_get_group_widths(h, widths) bind(C, name="qmcw_get_group_widths") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: widths(*)
      integer(c_int) :: st
    end function
    function qmcw_profile(h, enable) bind(C, name="qmcw_profile") result(st)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: enable
      integer(c_int) :: st
    end function
    function qmcw_profile_get(h, ms, count) bind(C, name="qmcw_profile_get") result(st)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: ms(*)
      integer(c_int64_t), intent(out) :: count(*)
      integer(c_int) :: st
    end function
  end interface
end module champ_amd_mod
