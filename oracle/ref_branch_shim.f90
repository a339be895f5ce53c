! ref_branch_shim.f90 -- bind(C) entry over the reference's own split_join
! (qmc/dmc/dmc_elec/split_join_mod.f90:3-219), compiled with the modules it
! uses from their own files, so that the oracle's branching and the engine's
! host split_join can be checked bit for bit on fixed weight vectors.
! Test infrastructure; compiled only into oracle/_ref.
subroutine ref_split_join(nw, cap, wt, tag, nout, nbr) bind(C, name="ref_split_join")
  use iso_c_binding
  use walker_type_mod
  use branch_mod, only: MWALK, pwt, wthist
  use age_mod, only: iage
  use velratio_mod, only: fratio
  use forcepar_mod, only: nforce
  use force_dmc_mod, only: nwprod
  use stats_mod, only: nbrnch
  use const_mod, only: ipr
  use split_join_mod
  implicit none
  integer(c_int32_t), value :: nw, cap
  real(c_double), intent(inout) :: wt(cap)       ! in: weights of nw walkers; out: of nout walkers
  integer(c_int32_t), intent(out) :: tag(cap)    ! out: 0-based index of the walker each one copies
  integer(c_int32_t), intent(out) :: nout, nbr   ! out: walkers after split_join, branching events
  type(walker_t), allocatable, target :: walkers(:)
  integer :: n, i
  MWALK = cap
  nforce = 1
  nwprod = 1
  ipr = -2
  nbrnch = 0
  if (allocated(pwt)) deallocate(pwt, wthist, iage, fratio)
  allocate(pwt(MWALK, nforce), wthist(MWALK, 0:nwprod - 1, nforce), iage(MWALK), fratio(MWALK, nforce))
  pwt = 1.d0; wthist = 1.d0; iage = 0; fratio = 1.d0
  allocate(walkers(cap))
  do i = 1, cap
    allocate(walkers(i)%x(1, 1))
    walkers(i)%x(1, 1) = dble(i - 1)             ! the walker's identity, copied with it
    walkers(i)%weight = 0.d0
    if (i <= nw) walkers(i)%weight = wt(i)
  enddo
  n = nw
  call split_join(walkers, n)
  nout = n
  nbr = nbrnch
  do i = 1, n
    wt(i) = walkers(i)%weight
    tag(i) = nint(walkers(i)%x(1, 1))
  enddo
  deallocate(walkers)
end subroutine

! the rannyu stream of this library (split_join draws rannyu(0))
subroutine ref_branch_setrn(iseed) bind(C, name="ref_branch_setrn")
  use iso_c_binding
  integer(c_int32_t), intent(in) :: iseed(4)
  integer :: is(4)
  is = iseed
  call setrn(is)
end subroutine

subroutine ref_branch_savern(iseed) bind(C, name="ref_branch_savern")
  use iso_c_binding
  integer(c_int32_t), intent(out) :: iseed(4)
  integer :: is(4)
  call savern(is)
  iseed = is
end subroutine

function ref_branch_rannyu() result(r) bind(C, name="ref_branch_rannyu")
  use iso_c_binding
  real(c_double) :: r
  real*8, external :: rannyu
  r = rannyu(0)
end function
