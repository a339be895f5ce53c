! qmcw_driver.f90 -- a Fortran program driving the engine through the
! iso_c_binding module (include/champ_amd_mod.f90), as CHAMP's Fortran
! drivers would: it fills a qmcw_system from a stream file, creates a handle,
! places the walkers, evaluates them (qmcw_eval_full), runs VMC steps
! (qmcw_vmc_run with the mover chosen by qmcw_set_vmc_mover) and writes the
! results to a stream file.
!
! usage: qmcw_driver <system file> <run file> <output file>
!   system file: the members of struct qmcw_system in declaration order;
!     int32 and double scalars as they are, each array as an int64 count
!     followed by its elements (count -1: a NULL pointer)
!   run file: int32 nwalk, nelec, nstep, imetro; double delta, fbias;
!     double x(3, nelec, nwalk); int32 seeds(4, nwalk)
!   output file: int32 status of each call; double logpsid, sgn, psij, eloc,
!     epot (nwalk each), grad(3, nelec, nwalk), lap(nelec, nwalk); after the
!     VMC run: double eloc trace(nwalk, nstep), int64 naccept(nwalk), double
!     x(3, nelec, nwalk), int32 seeds(4, nwalk)
module qmcw_driver_io
  use iso_c_binding
  implicit none
  type ibuf
    integer(c_int32_t), allocatable :: v(:)
  end type
  type dbuf
    real(c_double), allocatable :: v(:)
  end type
  type(ibuf), target, save :: ib(128)
  type(dbuf), target, save :: db(128)
  integer, save :: nib = 0, ndb = 0
contains
  function rd_i32(u) result(v)
    integer, intent(in) :: u
    integer(c_int32_t) :: v
    read(u) v
  end function
  function rd_f64(u) result(v)
    integer, intent(in) :: u
    real(c_double) :: v
    read(u) v
  end function
  function rd_iptr(u) result(p)
    integer, intent(in) :: u
    type(c_ptr) :: p
    integer(c_int64_t) :: n
    read(u) n
    if (n < 0) then
      p = c_null_ptr
      return
    end if
    nib = nib + 1
    if (nib > size(ib)) stop 'qmcw_driver: too many integer arrays'
    allocate(ib(nib)%v(max(n, 1_c_int64_t)))
    ib(nib)%v = 0
    if (n > 0) read(u) ib(nib)%v(1:n)
    p = c_loc(ib(nib)%v(1))
  end function
  function rd_dptr(u) result(p)
    integer, intent(in) :: u
    type(c_ptr) :: p
    integer(c_int64_t) :: n
    read(u) n
    if (n < 0) then
      p = c_null_ptr
      return
    end if
    ndb = ndb + 1
    if (ndb > size(db)) stop 'qmcw_driver: too many real arrays'
    allocate(db(ndb)%v(max(n, 1_c_int64_t)))
    db(ndb)%v = 0
    if (n > 0) read(u) db(ndb)%v(1:n)
    p = c_loc(db(ndb)%v(1))
  end function
  subroutine read_system(path, s)
    use champ_amd_mod, only: qmcw_system
    character(len=*), intent(in) :: path
    type(qmcw_system), intent(out) :: s
    integer :: u
    open(newunit=u, file=path, access='stream', form='unformatted', status='old', action='read')
  s%nelec = rd_i32(u)
  s%nup = rd_i32(u)
  s%ncent = rd_i32(u)
  s%nctype = rd_i32(u)
  s%cent = rd_dptr(u)
  s%iwctype = rd_iptr(u)
  s%znuc = rd_dptr(u)
  s%basis_type = rd_i32(u)
  s%nrad = rd_i32(u)
  s%rad_n = rd_iptr(u)
  s%rad_prim_off = rd_iptr(u)
  s%prim_zeta = rd_dptr(u)
  s%prim_coef = rd_dptr(u)
  s%ctype_bas_off = rd_iptr(u)
  s%bas_rad = rd_iptr(u)
  s%bas_l = rd_iptr(u)
  s%bas_m = rd_iptr(u)
  s%nbasis = rd_i32(u)
  s%norb = rd_i32(u)
  s%coef = rd_dptr(u)
  s%ndet = rd_i32(u)
  s%det_up = rd_iptr(u)
  s%det_dn = rd_iptr(u)
  s%ncsf = rd_i32(u)
  s%csf_coef = rd_dptr(u)
  s%csf_det_off = rd_iptr(u)
  s%csf_det = rd_iptr(u)
  s%csf_det_coef = rd_dptr(u)
  s%isc = rd_i32(u)
  s%nspin2b = rd_i32(u)
  s%nocuspb = rd_i32(u)
  s%norda = rd_i32(u)
  s%nordb = rd_i32(u)
  s%nordc = rd_i32(u)
  s%scalek = rd_f64(u)
  s%a4 = rd_dptr(u)
  s%b = rd_dptr(u)
  s%c = rd_dptr(u)
  s%nloc = rd_i32(u)
  s%npotd = rd_iptr(u)
  s%lpotp1 = rd_iptr(u)
  s%ecp_off = rd_iptr(u)
  s%ecp_coef = rd_dptr(u)
  s%ecp_power = rd_iptr(u)
  s%ecp_expo = rd_dptr(u)
  s%nquad = rd_i32(u)
  s%nparma_tot = rd_i32(u)
  s%jas_a_type = rd_iptr(u)
  s%jas_a_idx = rd_iptr(u)
  s%nparmb_tot = rd_i32(u)
  s%jas_b_spin = rd_iptr(u)
  s%jas_b_idx = rd_iptr(u)
  s%nparmc_tot = rd_i32(u)
  s%jas_c_type = rd_iptr(u)
  s%jas_c_idx = rd_iptr(u)
  s%nparmcsf = rd_i32(u)
  s%csf_parm = rd_iptr(u)
  s%nparmorb = rd_i32(u)
  s%orb_parm_i = rd_iptr(u)
  s%orb_parm_j = rd_iptr(u)
  s%orb_parm_rev = rd_iptr(u)
  s%hb = rd_f64(u)
  s%rad_num = rd_iptr(u)
  s%nnum = rd_i32(u)
  s%num_igrid = rd_iptr(u)
  s%num_nr = rd_iptr(u)
  s%num_h = rd_dptr(u)
  s%num_r0 = rd_dptr(u)
  s%num_off = rd_iptr(u)
  s%num_rwf = rd_dptr(u)
  s%num_d2 = rd_dptr(u)
  s%num_ce = rd_dptr(u)
  s%num_ae = rd_dptr(u)
    close(u)
  end subroutine
end module qmcw_driver_io

program qmcw_driver
  use iso_c_binding
  use champ_amd_mod
  use qmcw_driver_io
  implicit none
  type(qmcw_system), target :: s
  type(c_ptr) :: h
  character(len=4096) :: fsys, frun, fout
  integer :: u, nst
  integer(c_int) :: st(16)
  integer(c_int32_t) :: nwalk, nelec, nstep, imetro
  real(c_double) :: delta, fbias
  real(c_double), allocatable :: x(:), logpsid(:), sgn(:), psij(:), eloc(:), epot(:), grad(:), lap(:)
  real(c_double), allocatable, target :: trace(:)
  integer(c_int64_t), allocatable, target :: nacc(:)
  integer(c_int32_t), allocatable :: seeds(:)

  if (command_argument_count() /= 3) stop 'usage: qmcw_driver <system file> <run file> <output file>'
  call get_command_argument(1, fsys)
  call get_command_argument(2, frun)
  call get_command_argument(3, fout)
  call read_system(trim(fsys), s)
  if (qmcw_sizeof_system() /= c_sizeof(s)) stop 'qmcw_driver: qmcw_system layout differs from the library'

  open(newunit=u, file=trim(frun), access='stream', form='unformatted', status='old', action='read')
  read(u) nwalk, nelec, nstep, imetro
  read(u) delta, fbias
  allocate(x(3 * nelec * nwalk), seeds(4 * nwalk))
  read(u) x
  read(u) seeds
  close(u)
  allocate(logpsid(nwalk), sgn(nwalk), psij(nwalk), eloc(nwalk), epot(nwalk))
  allocate(grad(3 * nelec * nwalk), lap(nelec * nwalk))
  allocate(trace(nwalk * max(nstep, 1)), nacc(nwalk))

  st = 0
  nst = 0
  st(1) = qmcw_create(s, nwalk, QMCW_FP64, h)
  if (st(1) == QMCW_OK) then
    st(2) = qmcw_set_walkers(h, x)
    st(3) = qmcw_eval_full(h, logpsid, sgn, psij, grad, lap, eloc, epot)
    st(4) = qmcw_set_rng(h, seeds)
    st(5) = qmcw_set_vmc_mover(h, imetro, 5.0_c_double, 1.0_c_double)
    st(6) = qmcw_vmc_run(h, nstep, delta, fbias, c_loc(trace), c_loc(nacc))
    st(7) = qmcw_get_walkers(h, x)
    st(8) = qmcw_get_rng(h, seeds)
    st(9) = qmcw_destroy(h)
  end if

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) st
  write(u) logpsid, sgn, psij, eloc, epot, grad, lap
  write(u) trace(1:nwalk * nstep), nacc, x, seeds
  close(u)
  if (any(st /= QMCW_OK)) stop 1
end program qmcw_driver
