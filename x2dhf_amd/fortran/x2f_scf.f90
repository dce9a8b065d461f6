! SCF driver of one finite-difference HF job, with every array operation on
! the device through the x2gpu C ABI.  The routines keep the roles and the
! control flow of the reference's Fortran driver:
!   prep_scf        prepSCF                     doSCF.F90:24-174
!   scf_iteration   one pass of scf's loop      doSCF.F90:342-788
!   orbital_step    one orbital of that pass    doSCF.F90:368-533
!   relax_pots      coulExchSOR/MCSOR           relaxPots.F90:8-114,196-322
!   relax_orb       orbSOR/orbMCSOR             relaxOrbs.F90:16-241
!   fock_hf         fockHF                      fock.F90:57-184
!   ea_hf           EaHF                        lagrangeMultipliers.F90:21-107
!   norm / ortho    norm / ortho                normOrtho.F90:14-59,103-207
!   pot_asympt      potAsympt                   scfUtils.F90:179-222
!   mpole_moments   mpoleMoments/coulMom/exchMom scfUtils.F90:446-850
!   etotal_hf       etotalHF                    totalEnergy.F90:13-245
!   scf_finish      the tail of scf             doSCF.F90:793-858
! Closed shells only: the off-diagonal Lagrange multipliers of such cases are
! zero in the reference (offDiagLM is false for every pair, EabHF does
! nothing), so EabHF has no counterpart here.
!
! Two schedules:
!   schedule 0 ("plain"): the reference's order -- orbitals in reverse order,
!     each with its own Coulomb/exchange potentials relaxed first;
!   schedule 1 ("batched"): per iteration all potentials, then all orbitals
!     of one or many jobs are relaxed in single fused launches; the same
!     equations, a Jacobi instead of Gauss-Seidel order across orbitals.
module x2f_scf
  use, intrinsic :: iso_c_binding
  use, intrinsic :: ieee_arithmetic, only: ieee_is_nan
  use x2f_capi
  use x2f_grid
  use x2f_case
  implicit none
  private
  public :: scf_job, scf_opts, job_init, job_free, prep_scf, prep_scf_restart, scf_iteration, scf_finish, &
       batched_iteration, MAXPOT, job_bytes, control_probe

  integer, parameter :: dp = c_double
  integer, parameter :: MAXPOT = MAXO*(MAXO+1) + MAXO   ! exchange slots + Coulomb

  type :: scf_opts
     integer :: maxscf = 1000, ienterm = 10, inoterm = 8, nscfextra = 3, ienterm4init = 1
     integer :: maxsororb(2) = [1, 10], maxsorpot(2) = [1, 10]
     integer :: mpole = 4, prec = 0, schedule = 0
     integer :: nscf2skip = 600, nenlast = 20, nnolast = 20, maxresets = 5
     real(dp) :: omega_orb = -1.0_dp, omega_pot = -1.0_dp, recalc_mm = 1.15_dp
     ! orbEnergyIncThld: an orbital-energy change above it is divergence
     ! (initVariables.F90:187-191)
     real(dp) :: orb_energy_inc_thld = 1.0e4_dp
     ! FP64 finish of a mixed-precision job: the last `polish` iterations
     ! relax with FP64 sweeps (the reference's update), so the job stops at
     ! the reference's FP64 fixed point; 0 = none
     integer :: polish = 0
     ! relaxation scheme of FP64 sweeps (x2g_relax_item method / ordering):
     ! orbitals and potentials 0 MCSOR, 1 SOR; ordering an X2G_ORDER_* code
     integer :: orb_method = 0, pot_method = 0, ordering = 0
     ! altsweeps (doSCF.F90:403-411): SOR runs backward over the ordering
     ! for the odd orbitals and their potentials
     integer :: alt_sweeps = 0
     ! the batched schedule's end-of-iteration control: 0 on the device
     ! (x2g_scf_control), 1 by the host rules below (the plain schedule always
     ! uses the host rules)
     integer :: host_control = 0
  end type scf_opts

  type :: lc_batch
     type(x2g_lincomb_op), allocatable :: ops(:)
     type(x2g_term), allocatable :: terms(:)
     integer :: nops = 0, nterms = 0
  end type lc_batch

  type :: scf_job
     type(case_t) :: c
     type(grid_t) :: g
     type(scf_opts) :: o
     type(c_ptr) :: ctx = c_null_ptr, gd = c_null_ptr
     logical :: own_ctx = .true.
     type(c_ptr) :: S(14)                                  ! device coefficient arrays
     type(c_ptr) :: orb(MAXO), coul(MAXO), cpot(MAXO), xpot(MAXO)
     type(c_ptr) :: exch(MAXO*(MAXO+1))
     type(c_ptr) :: f1o(MAXO), f2o(MAXO), lo(MAXO)        ! per-orbital fock1, fock2, lhs
     type(c_ptr) :: plhs(MAXPOT), prhs(MAXPOT)             ! per-potential lhs, rhs
     integer :: pidx(MAXO, 2*MAXO+1) = 0                   ! (orbital, k) -> potential index
     type(c_ptr) :: ones = c_null_ptr
     integer(c_size_t) :: n = 0
     integer :: ndev = 0
     type(c_ptr) :: devbufs(512)
     ! state
     real(dp) :: ee(MAXO,MAXO) = 0, eeold(MAXO) = 0, deltaee(MAXO) = 0, orbnorm(MAXO) = 1
     real(dp) :: cmulti(8,MAXO) = 0, exc(8,MAXO*(MAXO+1)) = 0
     integer :: ifixorb(MAXO) = 0, itouch(MAXO) = 0, iasympt = 0
     ! a multipole-moment refresh decided by convergence_control, with the
     ! orbitals touched in that iteration (done by the caller, several jobs
     ! in one launch in the batched schedule)
     logical :: mm_due = .false.
     integer :: mm_touch(MAXO) = 0
     logical :: lfixorb = .false., lfixorb4init = .false.
     real(dp) :: omega_orb = 0, omega_pot = 0
     ! control state of the SCF loop
     integer :: iscf = 0, nresets = 0, nresets4no = 0, norbenergythld = 0, norbnormthld = 0
     integer :: noenergydec = 0, nonormdec = 0
     logical :: lenergythld = .false., lnormthld = .false., done = .false.
     real(dp) :: eeconv = 10, eeconvprev = 1.e35_dp, dnmaxprev = 1.e35_dp
     real(dp) :: deltaeemax = 100, deltanormmax = 100
     integer :: polish_end = 0     ! no threshold stop before this iteration (FP64 finish)
     integer :: stopcode = 0       ! 1 energy, 2 norm, 3 iteration limit, 4 saturation, 5 resets,
                                   ! 6 divergence, 7 not a number
     integer(c_long_long) :: nsweep_orb = 0, nsweep_pot = 0
     ! device record written by the batched schedule: <psi|psi> before
     ! normalisation (1..MAXO) and orbital energies (MAXO+1..2*MAXO),
     ! followed by the job's device control record (x2g_scf_ctrl), which the
     ! batched schedule advances on the device; ctrl_dirty: the host state
     ! changed since the record was last written
     type(c_ptr) :: rec = c_null_ptr, ctrl = c_null_ptr
     logical :: ctrl_dirty = .true.
     ! results
     real(dp) :: etot = 0, vkt = 0, vnt = 0, ecoul = 0, eexch = 0, virial = 0
  end type scf_job

  real(dp), parameter :: eps0 = epsilon(1.0_dp)

contains

  ! ---------------------------------------------------------------------------
  ! set-up: grid on the device, starting functions, device buffers
  ! ctx_in: an existing x2gpu context to share (jobs relaxed together in one
  ! batched launch must share one), or c_null_ptr for a context of its own.
  integer function job_init(j, device, ctx_in) result(ierr)
    type(scf_job), intent(inout), target :: j
    integer, intent(in) :: device
    type(c_ptr), intent(in) :: ctx_in
    real(dp) :: coef(25)
    real(dp), allocatable, target :: vec(:)
    type(x2g_lcao) :: lc
    integer :: i, k, ip, ip2, ii, slot
    integer :: pdm_of(MAXPOT)
    integer(c_size_t) :: nb

    ierr = 0
    if (c_associated(ctx_in)) then
       j%ctx = ctx_in
       j%own_ctx = .false.
    else
       call x2check(x2g_ctx_create(int(device, c_int), j%ctx), 'x2g_ctx_create')
       j%own_ctx = .true.
    endif
    allocate(vec(2*j%g%nmu + 2*j%g%nni))
    call grid_coef(j%g, coef)
    call grid_vec(j%g, vec)
    call x2check(x2g_grid_create(j%ctx, int(j%g%nni, c_int), int(j%g%nmu, c_int), coef, j%g%suppl, vec, j%gd), &
         'x2g_grid_create')
    do k = 1, 14
       j%S(k) = x2g_grid_suppl(j%gd, int(k-1, c_int))
    enddo
    j%n = int(j%g%nni, c_size_t)*int(j%g%nmu, c_size_t)
    nb = 8_c_size_t*j%n
    j%ndev = 0
    do i = 1, j%c%norb
       j%orb(i) = dalloc(j); j%coul(i) = dalloc(j); j%cpot(i) = dalloc(j); j%xpot(i) = dalloc(j)
       j%f1o(i) = dalloc(j); j%f2o(i) = dalloc(j); j%lo(i) = dalloc(j)
    enddo
    do slot = 1, j%c%nslots
       j%exch(slot) = dalloc(j)
    enddo
    j%ones = dalloc(j)
    call x2check(x2g_fill(j%ctx, j%ones, j%n, 1.0_dp), 'x2g_fill')
    block
      type(x2g_scf_ctrl), target :: c0
      if (x2g_scf_ctrl_bytes() /= int(c_sizeof(c0), c_int) .or. X2G_CTRL_MAXO /= MAXO) &
           error stop 'x2g_scf_ctrl: the C library and the Fortran driver lay the record out differently'
      call x2check(x2g_malloc(j%ctx, int(16*MAXO, c_size_t) + c_sizeof(c0), j%rec), 'x2g_malloc')
      call x2check(x2g_fill(j%ctx, j%rec, int(2*MAXO, c_size_t), 0.0_dp), 'x2g_fill')
      j%ctrl = ptr_off(j%rec, 16*MAXO)
    end block
    j%ctrl_dirty = .true.
    j%ndev = j%ndev + 1
    j%devbufs(j%ndev) = j%rec
    ! potentials of the iteration, in the SCF's orbital order
    ip = 0
    do ii = 1, j%c%norb
       i = j%c%norb - ii + 1
       do k = 1, j%c%npot(i)
          ip = ip + 1
          j%pidx(i,k) = ip
          j%prhs(ip) = dalloc(j)
          ! lhs = f3 + dm^2 e (relaxPots.F90:87-95), fixed for the job: one
          ! array per distinct dm, shared by the potentials with that dm
          ! (their relaxations then read it from cache)
          j%plhs(ip) = c_null_ptr
          do ip2 = 1, ip - 1
             if (pdm_of(ip2) == j%c%pdm(i,k)) then
                j%plhs(ip) = j%plhs(ip2)
                exit
             endif
          enddo
          pdm_of(ip) = j%c%pdm(i,k)
          if (c_associated(j%plhs(ip))) cycle
          j%plhs(ip) = dalloc(j)
          if (j%c%pdm(i,k) == 0) then
             call lincomb0(j, j%plhs(ip), j%S(8), 1.0_dp)
          else
             call lincomb1(j, j%plhs(ip), j%S(8), 1.0_dp, dble(j%c%pdm(i,k)**2), j%S(4), c_null_ptr, c_null_ptr)
          endif
       enddo
    enddo

    ! starting orbitals and potentials, formed on the device (initHyd +
    ! initCoulomb + initExchange); the exchange start of the pair (i, i') is
    ! that of its first orbital i
    do i = 1, j%c%norb
       lc = x2g_lcao(int(j%c%m(i), c_int), int(j%c%n1(i), c_int), int(j%c%l1(i), c_int), &
            int(j%c%n2(i), c_int), int(j%c%l2(i), c_int), j%c%c1(i), j%c%za(i), j%c%c2(i), j%c%zb(i))
       call x2check(x2g_guess_orbital(j%gd, j%orb(i), lc), 'x2g_guess_orbital')
       call x2check(x2g_guess_coulomb(j%gd, j%coul(i), j%c%c1(i), j%c%c2(i), &
            merge(1_c_int, 0_c_int, nint(j%c%z1) /= 0 .and. nint(j%c%z2) /= 0)), 'x2g_guess_coulomb')
    enddo
    do i = 1, j%c%norb
       do ii = i, j%c%norb
          k = pair_index(i, ii)
          if (j%c%ilc(k) == 0) cycle
          call x2check(x2g_guess_exchange(j%gd, j%exch(j%c%slot(k)), j%c%c1(i), j%c%c2(i)), 'x2g_guess_exchange')
          if (j%c%ilc(k) == 2) call x2check(x2g_guess_exchange(j%gd, j%exch(j%c%slot2(k)), j%c%c1(i), &
               j%c%c2(i)), 'x2g_guess_exchange')
       enddo
    enddo
    deallocate(vec)

    ! relaxation parameters
    call omega_defaults(j%g%nni, j%g%nmu, j%c%z1, j%c%z2, j%omega_orb, j%omega_pot)
    if (j%o%omega_orb > 0.0_dp) j%omega_orb = j%o%omega_orb
    if (j%o%omega_pot > 0.0_dp) j%omega_pot = j%o%omega_pot
    ! a hydrogenic start keeps the orbitals frozen until the orbital energies
    ! change by less than 10^-ienterm4init (read_orbpot, inputInterface.F90:2273-2291)
    j%lfixorb4init = j%o%ienterm4init > 0
    j%lfixorb = j%lfixorb4init
    j%ifixorb = merge(1, 0, j%lfixorb4init)
    j%ee = 0
  end function job_init

  function dalloc(j) result(p)
    type(scf_job), intent(inout) :: j
    type(c_ptr) :: p
    call x2check(x2g_malloc(j%ctx, 8_c_size_t*j%n, p), 'x2g_malloc')
    call x2check(x2g_fill(j%ctx, p, j%n, 0.0_dp), 'x2g_fill')
    j%ndev = j%ndev + 1
    j%devbufs(j%ndev) = p
  end function dalloc

  integer(c_size_t) function job_bytes(j)
    type(scf_job), intent(in) :: j
    job_bytes = 8_c_size_t*j%n*int(j%ndev + 14, c_size_t)
  end function job_bytes

  subroutine job_free(j)
    type(scf_job), intent(inout) :: j
    integer :: i
    integer(c_int) :: rc
    if (.not. c_associated(j%ctx)) return
    do i = 1, j%ndev
       rc = x2g_free(j%ctx, j%devbufs(i))
    enddo
    j%ndev = 0
    if (c_associated(j%gd)) rc = x2g_grid_destroy(j%gd)
    if (j%own_ctx) rc = x2g_ctx_destroy(j%ctx)
    j%ctx = c_null_ptr
    j%gd = c_null_ptr
  end subroutine job_free

  ! ---------------------------------------------------------------------------
  ! small wrappers of x2g_lincomb / x2g_wdot
  subroutine lincomb0(j, out, base, s)
    type(scf_job), intent(inout) :: j
    type(c_ptr), intent(in) :: out, base
    real(dp), intent(in) :: s
    type(x2g_term) :: t(1)
    call x2check(x2g_lincomb(j%gd, out, base, s, t, 0_c_int), 'x2g_lincomb')
  end subroutine lincomb0

  subroutine lincomb1(j, out, base, s, c, a, b, cc)
    type(scf_job), intent(inout) :: j
    type(c_ptr), intent(in) :: out, base, a, b, cc
    real(dp), intent(in) :: s, c
    type(x2g_term) :: t(1)
    t(1) = x2g_term(c, a, b, cc)
    call x2check(x2g_lincomb(j%gd, out, base, s, t, 1_c_int), 'x2g_lincomb')
  end subroutine lincomb1

  ! ---------------------------------------------------------------------------
  ! fockHF (fock.F90:57-184): fock1 = f0 + eps f1 + m^2 e + f2 * sum_j coo_j J_j,
  ! fock2 = f2 * sum_j psi_j (gec K_ij [+ gec' K'_ij] + eps_ji f4) [+ self term].
  ! The Coulomb and exchange sums are kept per orbital for EaHF.
  ! fock_terms returns the Coulomb-sum terms (tc), the exchange-sum terms (tx)
  ! and the fock1 terms other than f2 * Coulomb sum (t1).
  subroutine fock_terms(j, i, tc, ntc, tx, ntx, t1, nt1)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    type(x2g_term), intent(out) :: tc(:), tx(:), t1(:)
    integer, intent(out) :: ntc, ntx, nt1
    integer :: i1, k
    real(dp) :: coo
    nt1 = 1
    t1(1) = x2g_term(j%ee(i,i), j%S(6), c_null_ptr, c_null_ptr)
    if (j%c%m(i) /= 0) then
       nt1 = nt1 + 1; t1(nt1) = x2g_term(dble(j%c%m(i)**2), j%S(4), c_null_ptr, c_null_ptr)
    endif
    ntc = 0; ntx = 0
    if (j%c%nel == 1) return
    do i1 = 1, j%c%norb
       coo = j%c%occ(i1)
       if (i1 == i) coo = coo - 1.0_dp
       ntc = ntc + 1; tc(ntc) = x2g_term(coo, j%coul(i1), c_null_ptr, c_null_ptr)
    enddo
    do i1 = 1, j%c%norb
       k = j%c%k2(i, i1)
       if (i1 /= i) then
          ntx = ntx + 1; tx(ntx) = x2g_term(j%c%gec(i,i1), j%exch(j%c%slot(k)), j%orb(i1), c_null_ptr)
          if (j%c%ilc(k) == 2) then
             ntx = ntx + 1; tx(ntx) = x2g_term(j%c%gec2(i,i1), j%exch(j%c%slot2(k)), j%orb(i1), c_null_ptr)
          endif
          if (abs(j%ee(i1,i)) > eps0) then
             ntx = ntx + 1; tx(ntx) = x2g_term(j%ee(i1,i), j%S(9), j%orb(i1), c_null_ptr)
          endif
       else
          if (j%c%m(i) > 0 .and. k > 0) then
             if (j%c%ilc(k) > 0) then
                ntx = ntx + 1; tx(ntx) = x2g_term(j%c%gec(i,i), j%exch(j%c%slot(k)), j%orb(i), c_null_ptr)
             endif
          endif
       endif
    enddo
  end subroutine fock_terms

  subroutine fock_hf(j, i, fock1, fock2)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    type(c_ptr), intent(in) :: fock1, fock2
    type(x2g_term) :: t(4), tc(MAXO), tx(4*MAXO)
    integer :: n, ntc, ntx
    call fock_terms(j, i, tc, ntc, tx, ntx, t, n)
    if (j%c%nel == 1) then
       call x2check(x2g_lincomb(j%gd, fock1, j%S(5), 1.0_dp, t, int(n, c_int)), 'fock1')
       call x2check(x2g_fill(j%ctx, fock2, j%n, 0.0_dp), 'fock2')
       return
    endif
    call x2check(x2g_lincomb(j%gd, j%cpot(i), c_null_ptr, 0.0_dp, tc, int(ntc, c_int)), 'fock coulomb')
    call x2check(x2g_lincomb(j%gd, j%xpot(i), c_null_ptr, 0.0_dp, tx, int(ntx, c_int)), 'fock exchange')
    n = n + 1; t(n) = x2g_term(1.0_dp, j%S(7), j%cpot(i), c_null_ptr)
    call x2check(x2g_lincomb(j%gd, fock1, j%S(5), 1.0_dp, t, int(n, c_int)), 'fock1')
    t(1) = x2g_term(1.0_dp, j%S(7), j%xpot(i), c_null_ptr)
    call x2check(x2g_lincomb(j%gd, fock2, c_null_ptr, 0.0_dp, t, 1_c_int), 'fock2')
  end subroutine fock_hf

  ! ---------------------------------------------------------------------------
  ! Collected linear combinations, launched together by x2g_lincomb_batch.
  subroutine lc_add(b, out, base, s, t, nt, gidx)
    type(lc_batch), intent(inout) :: b
    type(c_ptr), intent(in) :: out, base
    real(dp), intent(in) :: s
    type(x2g_term), intent(in) :: t(:)
    integer, intent(in) :: nt, gidx
    type(x2g_lincomb_op), allocatable :: o2(:)
    type(x2g_term), allocatable :: t2(:)
    if (.not. allocated(b%ops)) allocate(b%ops(64), b%terms(256))
    if (b%nops == size(b%ops)) then
       allocate(o2(2*size(b%ops))); o2(1:b%nops) = b%ops(1:b%nops)
       call move_alloc(o2, b%ops)
    endif
    if (b%nterms + nt > size(b%terms)) then
       allocate(t2(2*(b%nterms + nt))); t2(1:b%nterms) = b%terms(1:b%nterms)
       call move_alloc(t2, b%terms)
    endif
    b%nops = b%nops + 1
    b%ops(b%nops) = x2g_lincomb_op(out, base, s, int(b%nterms, c_int), int(nt, c_int), int(gidx, c_int))
    if (nt > 0) b%terms(b%nterms+1:b%nterms+nt) = t(1:nt)
    b%nterms = b%nterms + nt
  end subroutine lc_add

  subroutine lc_run(b, ctx, grids)
    type(lc_batch), intent(inout) :: b
    type(c_ptr), intent(in) :: ctx
    type(c_ptr), intent(in) :: grids(:)
    if (b%nops == 0) return
    call x2check(x2g_lincomb_batch(ctx, grids, b%ops, int(b%nops, c_int), b%terms, int(b%nterms, c_int)), &
         'x2g_lincomb_batch')
    b%nops = 0
    b%nterms = 0
  end subroutine lc_run

  ! c_ptr advanced by nbytes
  function ptr_off(p, nbytes) result(q)
    type(c_ptr), intent(in) :: p
    integer, intent(in) :: nbytes
    type(c_ptr) :: q
    q = transfer(transfer(p, 0_c_intptr_t) + int(nbytes, c_intptr_t), q)
  end function ptr_off

  ! ---------------------------------------------------------------------------
  ! relaxation items of the potentials of orbital i (relaxPots): rhs = psi1 psi2 g
  subroutine pot_items(j, i, items, n, gidx, lb)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i, gidx
    type(x2g_relax_item), intent(inout) :: items(:)
    integer, intent(inout) :: n
    type(lc_batch), intent(inout), optional :: lb   ! collect the rhs products instead of launching them
    integer :: k, ip
    type(c_ptr) :: fun
    type(x2g_term) :: t(1)
    do k = 1, j%c%npot(i)
       ip = j%pidx(i,k)
       if (present(lb) .and. j%o%prec == 1) then
          ! the FP64 residual of the correction forms psi1 psi2 g itself (below)
       else if (present(lb)) then
          t(1) = x2g_term(1.0_dp, j%orb(j%c%pin1(i,k)), j%orb(j%c%pin2(i,k)), j%S(10))
          call lc_add(lb, j%prhs(ip), c_null_ptr, 0.0_dp, t, 1, gidx)
       else
          call lincomb1(j, j%prhs(ip), c_null_ptr, 0.0_dp, 1.0_dp, j%orb(j%c%pin1(i,k)), j%orb(j%c%pin2(i,k)), &
               j%S(10))
       endif
       if (j%c%pslot(i,k) == 0) then
          fun = j%coul(i)
       else
          fun = j%exch(j%c%pslot(i,k))
       endif
       n = n + 1
       items(n) = x2g_relax_item(fun, j%plhs(ip), j%prhs(ip), c_null_ptr, j%omega_pot, &
            int(j%c%psym(i,k), c_int), 1_c_int, int(j%o%prec, c_int), int(gidx, c_int))
       call set_scheme(j, items(n), j%o%pot_method, i)
       if (present(lb) .and. j%o%prec == 1) then
          items(n)%rhs_f1_d = j%orb(j%c%pin1(i,k))
          items(n)%rhs_f2_d = j%orb(j%c%pin2(i,k))
          items(n)%rhs_f3_d = j%S(10)
       endif
    enddo
    j%nsweep_pot = j%nsweep_pot + int(j%c%npot(i), c_long_long)*j%o%maxsorpot(1)*j%o%maxsorpot(2)
  end subroutine pot_items

  ! the relaxation item of orbital i (orbMCSOR): lhs = fock1 + diag, rhs = fock2
  subroutine orb_item(j, i, items, n, gidx)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i, gidx
    type(x2g_relax_item), intent(inout) :: items(:)
    integer, intent(inout) :: n
    call fock_hf(j, i, j%f1o(i), j%f2o(i))
    call lincomb1(j, j%lo(i), j%f1o(i), 1.0_dp, j%g%diag, j%ones, c_null_ptr, c_null_ptr)
    n = n + 1
    items(n) = x2g_relax_item(j%orb(i), j%lo(i), j%f2o(i), j%f1o(i), j%omega_orb, &
         int(j%c%isym(i), c_int), 0_c_int, int(j%o%prec, c_int), int(gidx, c_int))
    call set_scheme(j, items(n), j%o%orb_method, i)
    j%nsweep_orb = j%nsweep_orb + int(j%o%maxsororb(1)*j%o%maxsororb(2), c_long_long)
  end subroutine orb_item

  ! The relaxation scheme of an FP64 item of orbital iorb (the FP32
  ! correction sweeps are col-wise MCSOR whatever the job asks for).  With
  ! altsweeps the SOR of an odd orbital and of its potentials walks the
  ! ordering backwards (isstep = -1, doSCF.F90:403-411); MCSOR ignores it.
  subroutine set_scheme(j, it, method, iorb)
    type(scf_job), intent(in) :: j
    type(x2g_relax_item), intent(inout) :: it
    integer, intent(in) :: method, iorb
    if (j%o%prec /= 0) return
    it%method = int(method, c_int)
    if (method == 1 .and. j%o%alt_sweeps /= 0 .and. mod(iorb, 2) == 1) it%method = 2_c_int
    it%ordering = int(j%o%ordering, c_int)
  end subroutine set_scheme

  subroutine relax_pots(j, i)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    type(x2g_relax_item) :: items(2*MAXO+1)
    type(c_ptr) :: grids(1)
    integer :: n
    n = 0
    call pot_items(j, i, items, n, 0)
    grids(1) = j%gd
    call x2check(x2g_relax_batch(j%ctx, grids, items, int(n, c_int), int(j%o%maxsorpot(1), c_int), &
         int(j%o%maxsorpot(2), c_int)), 'relax potentials')
  end subroutine relax_pots

  subroutine relax_orb(j, i)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    type(x2g_relax_item) :: items(1)
    type(c_ptr) :: grids(1)
    integer :: n
    if (j%ifixorb(i) == 1) return
    n = 0
    call orb_item(j, i, items, n, 0)
    grids(1) = j%gd
    call x2check(x2g_relax_batch(j%ctx, grids, items, 1_c_int, int(j%o%maxsororb(1), c_int), &
         int(j%o%maxsororb(2), c_int)), 'relax orbital')
  end subroutine relax_orb

  ! ---------------------------------------------------------------------------
  ! norm (normOrtho.F90:14-59)
  subroutine norm(j, i)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    type(x2g_dot) :: d(1)
    real(dp) :: x(1)
    d(1) = x2g_dot(j%S(14), j%orb(i), j%orb(i), j%S(9))
    call x2check(x2g_wdot(j%gd, d, 1_c_int, x), 'norm')
    j%orbnorm(i) = x(1)
    call lincomb0(j, j%orb(i), j%orb(i), 1.0_dp/sqrt(x(1)))
  end subroutine norm

  ! ortho (normOrtho.F90:103-190): Gram-Schmidt against the higher orbitals of
  ! the same symmetry, then the reference's rescaling
  subroutine ortho(j, i1)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i1
    type(x2g_dot) :: d(MAXO)
    type(x2g_term) :: t(MAXO)
    real(dp) :: ovla(MAXO), ano
    integer :: jor, i2, i3, k
    if (j%ifixorb(i1) /= 0) return
    jor = 0
    do i2 = 1, j%c%norb
       i3 = j%c%norb - i2 + 1
       if (i3 <= i1) cycle
       if (j%c%m(i3) /= j%c%m(i1)) cycle
       if (j%c%ige(i3) /= j%c%ige(i1)) cycle
       jor = jor + 1
       d(jor) = x2g_dot(j%S(14), j%orb(i1), j%orb(i3), j%S(9))
       t(jor)%a_d = j%orb(i3)
    enddo
    if (jor == 0) return
    call x2check(x2g_wdot(j%gd, d, int(jor, c_int), ovla), 'ortho overlaps')
    ano = 0.0_dp
    do k = 1, jor
       ano = ano + ovla(k)*ovla(k)
       t(k)%c = -ovla(k)
       t(k)%b_d = c_null_ptr
       t(k)%cc_d = c_null_ptr
    enddo
    call x2check(x2g_lincomb(j%gd, j%orb(i1), j%orb(i1), 1.0_dp, t, int(jor, c_int)), 'ortho')
    ano = sqrt(1.0_dp + ano)/(1.0_dp - ano*ano)
    call lincomb0(j, j%orb(i1), j%orb(i1), ano)
  end subroutine ortho

  subroutine setci(j, i)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    integer :: k
    if (.not. j%c%homonuclear) return
    if (i == 0) then
       do k = 1, j%c%norb
          call x2check(x2g_setci(j%gd, j%orb(k), int(j%c%ihomo(k), c_int)), 'setci')
       enddo
    else
       call x2check(x2g_setci(j%gd, j%orb(i), int(j%c%ihomo(i), c_int)), 'setci')
    endif
  end subroutine setci

  ! EaHF (lagrangeMultipliers.F90:21-107) with the Coulomb and exchange sums of
  ! the last fock_hf of orbital i
  subroutine ea_hf(j, i)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    real(dp) :: one(3), two(2)
    type(x2g_dot) :: d(2)
    call x2check(x2g_oneel(j%gd, j%orb(i), j%orb(i), int(j%c%isym(i), c_int), int(j%c%m(i), c_int), one), 'oneel')
    d(1) = x2g_dot(j%S(14), j%orb(i), j%cpot(i), j%orb(i))
    d(2) = x2g_dot(j%S(14), j%orb(i), j%xpot(i), c_null_ptr)
    call x2check(x2g_wdot(j%gd, d, 2_c_int, two), 'EaHF')
    j%ee(i,i) = one(1) + (two(1) - two(2))
  end subroutine ea_hf

  ! ---------------------------------------------------------------------------
  ! mpoleMoments: Coulomb moments of touched orbitals (odd ones vanish by
  ! symmetry for homonuclear molecules) and the moments of every exchange pair,
  ! all in one batched reduction
  subroutine mpole_moments(j, touch)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: touch(:)
    type(x2g_moment_item), allocatable :: it(:)
    real(dp), allocatable :: out(:)
    integer, allocatable :: who(:)
    type(c_ptr) :: grids(1)
    integer :: n
    if (j%c%nel == 1) return
    allocate(it(j%c%norb + j%c%nmm), who(j%c%norb + j%c%nmm))
    n = 0
    call mm_items(j, touch, 0, it, who, n)
    if (n == 0) return
    allocate(out(8*n))
    grids(1) = j%gd
    call x2check(x2g_moments_batch(j%ctx, grids, it, int(n, c_int), int(j%o%mpole, c_int), out), 'mpoleMoments')
    call mm_store(j, who, out, n)
  end subroutine mpole_moments

  ! the moment items of job j (coulMom / exchMom): the densities of the
  ! touched orbitals and the exchange pairs with their Delta m, on grid gidx
  subroutine mm_items(j, touch, gidx, it, who, n)
    type(scf_job), intent(in) :: j
    integer, intent(in) :: touch(:), gidx
    type(x2g_moment_item), intent(inout) :: it(:)
    integer, intent(inout) :: who(:), n
    integer :: i, k
    do i = 1, j%c%norb
       if (touch(i) == 0) cycle
       n = n + 1
       it(n) = x2g_moment_item(j%orb(i), j%orb(i), 0_c_int, int(gidx, c_int))
       who(n) = i
    enddo
    do k = 1, j%c%nmm
       n = n + 1
       it(n) = x2g_moment_item(j%orb(j%c%mm1(k)), j%orb(j%c%mm2(k)), int(j%c%mmdel(k), c_int), int(gidx, c_int))
       who(n) = -k
    enddo
  end subroutine mm_items

  ! moments out(8*(k-1)+1..) of the items who(1..n) of job j into its state
  subroutine mm_store(j, who, out, n)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: who(:), n
    real(dp), intent(in) :: out(:)
    integer :: i, k, mp
    mp = j%o%mpole
    do k = 1, n
       if (who(k) > 0) then
          i = who(k)
          j%cmulti(1:mp, i) = out(8*(k-1)+1:8*(k-1)+mp)
          if (j%c%homonuclear) then
             j%cmulti(1, i) = 0.0_dp; j%cmulti(3, i) = 0.0_dp
             j%cmulti(5, i) = 0.0_dp; j%cmulti(7, i) = 0.0_dp
          endif
       else
          j%exc(1:mp, j%c%mmslot(-who(k))) = out(8*(k-1)+1:8*(k-1)+mp)
       endif
    enddo
  end subroutine mm_store

  ! the moment refreshes that convergence_control scheduled for jobs
  ! js(idx(q)), q = 1..nj, in one launch and one transfer
  subroutine mm_refresh(js, idx, nj)
    type(scf_job), intent(inout), target :: js(:)
    integer, intent(in) :: idx(:), nj
    type(x2g_moment_item), allocatable :: it(:)
    type(c_ptr), allocatable :: grids(:)
    integer, allocatable :: who(:), first(:)
    real(dp), allocatable :: out(:)
    integer :: q, h, n, nmax, mp
    nmax = 0; mp = 2
    allocate(grids(nj), first(nj + 1))
    do q = 1, nj
       h = idx(q)
       grids(q) = js(h)%gd
       if (.not. js(h)%mm_due .or. js(h)%c%nel == 1) cycle
       nmax = nmax + js(h)%c%norb + js(h)%c%nmm
       mp = max(mp, js(h)%o%mpole)
    enddo
    if (nmax > 0) then
       allocate(it(nmax), who(nmax), out(8*nmax))
       n = 0
       do q = 1, nj
          h = idx(q)
          first(q) = n
          if (.not. js(h)%mm_due .or. js(h)%c%nel == 1) cycle
          call mm_items(js(h), js(h)%mm_touch, q-1, it, who, n)
       enddo
       first(nj + 1) = n
       if (n > 0) then
          call x2check(x2g_moments_batch(js(idx(1))%ctx, grids, it, int(n, c_int), int(mp, c_int), out), &
               'batched mpoleMoments')
          do q = 1, nj
             h = idx(q)
             if (first(q + 1) > first(q)) &
                  call mm_store(js(h), who(first(q)+1:first(q+1)), out(8*first(q)+1:8*first(q+1)), &
                  first(q+1) - first(q))
          enddo
       endif
    endif
    do q = 1, nj
       js(idx(q))%mm_due = .false.
    enddo
  end subroutine mm_refresh

  ! potAsympt: the multipole-tail items of the potentials of orbital i1; a
  ! slot already queued (done) is not queued again
  subroutine pot_tail_items(j, i1, ti, nt, gidx, done)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i1, gidx
    type(x2g_tail_item), intent(inout) :: ti(:)
    integer, intent(inout) :: nt
    logical, intent(inout) :: done(:)
    integer :: i2, k, idel
    if (j%itouch(i1) == 1) then
       nt = nt + 1
       ti(nt) = x2g_tail_item(j%coul(i1), 0_c_int, 0_c_int, int(gidx, c_int), j%cmulti(:, i1))
    endif
    do i2 = 1, j%c%norb
       if (j%itouch(i1) == 0 .and. j%itouch(i2) == 0) cycle
       k = j%c%k2(i1, i2)
       if (k == 0) cycle
       if (j%c%ilc(k) == 0) cycle
       idel = abs(j%c%m(i1) - j%c%m(i2))
       if (i1 == i2) idel = 2*j%c%m(i1)
       if (.not. done(j%c%slot(k))) then
          nt = nt + 1
          ti(nt) = x2g_tail_item(j%exch(j%c%slot(k)), 1_c_int, int(idel, c_int), int(gidx, c_int), &
               j%exc(:, j%c%slot(k)))
          done(j%c%slot(k)) = .true.
       endif
       if (j%c%ilc(k) == 2) then
          if (.not. done(j%c%slot2(k))) then
             idel = j%c%m(i2) + j%c%m(i1)
             nt = nt + 1
             ti(nt) = x2g_tail_item(j%exch(j%c%slot2(k)), 1_c_int, int(idel, c_int), int(gidx, c_int), &
                  j%exc(:, j%c%slot2(k)))
             done(j%c%slot2(k)) = .true.
          endif
       endif
    enddo
  end subroutine pot_tail_items

  ! potAsympt of orbital i1 in one launch
  subroutine pot_asympt(j, i1)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i1
    type(x2g_tail_item) :: ti(1 + 2*MAXO)
    logical :: done(MAXO*(MAXO+1))
    type(c_ptr) :: grids(1)
    integer :: nt
    nt = 0
    done = .false.
    call pot_tail_items(j, i1, ti, nt, 0, done)
    grids(1) = j%gd
    call x2check(x2g_pot_tail_batch(j%ctx, grids, ti, int(nt, c_int), int(j%o%mpole, c_int)), 'potAsympt')
  end subroutine pot_asympt

  ! ---------------------------------------------------------------------------
  ! etotalHF (totalEnergy.F90:13-245)
  subroutine etotal_hf(j)
    type(scf_job), intent(inout) :: j
    type(x2g_dot) :: d(MAXO*MAXO*2 + MAXO)
    real(dp) :: v(MAXO*MAXO*2 + MAXO), one(3), vk(MAXO), vn(MAXO), oneel(MAXO)
    real(dp) :: woneel, wdcoul, wndc, wex1, wex2, evt, epott
    integer :: i1, i2, nd, k, idx
    ! one-electron parts
    do i1 = 1, j%c%norb
       call x2check(x2g_oneel(j%gd, j%orb(i1), j%orb(i1), int(j%c%isym(i1), c_int), int(j%c%m(i1), c_int), one), &
            'oneelii')
       oneel(i1) = one(1); vk(i1) = one(2); vn(i1) = one(3)
    enddo
    ! all two-electron quadratures in one batch, combined below in the reference's order
    nd = 0
    do i1 = 1, j%c%norb
       nd = nd + 1; d(nd) = x2g_dot(j%S(14), j%orb(i1), j%coul(i1), j%orb(i1))            ! J_ii
       if (j%c%m(i1) > 0) then
          k = j%c%k2(i1, i1)
          nd = nd + 1; d(nd) = x2g_dot(j%S(14), j%orb(i1), j%exch(j%c%slot(k)), j%orb(i1)) ! K_ii (dm = 2m)
       endif
       do i2 = i1+1, j%c%norb
          k = j%c%k2(i1, i2)
          nd = nd + 1; d(nd) = x2g_dot(j%S(14), j%orb(i1), j%coul(i2), j%orb(i1))        ! coulij(i2,i1)
          nd = nd + 1; d(nd) = x2g_dot(j%S(14), j%orb(i1), j%exch(j%c%slot(k)), j%orb(i2))
          if (j%c%ilc(k) == 2) then
             nd = nd + 1; d(nd) = x2g_dot(j%S(14), j%orb(i1), j%exch(j%c%slot2(k)), j%orb(i2))
          endif
       enddo
    enddo
    call x2check(x2g_wdot(j%gd, d, int(nd, c_int), v), 'etotal')
    woneel = 0; wdcoul = 0; wndc = 0; wex1 = 0; wex2 = 0
    j%vkt = 0; j%vnt = 0
    idx = 0
    do i1 = 1, j%c%norb
       woneel = woneel + j%c%occ(i1)*oneel(i1)
       j%vkt = j%vkt + j%c%occ(i1)*vk(i1)
       j%vnt = j%vnt + j%c%occ(i1)*vn(i1)
       idx = idx + 1
       if (j%c%occ(i1) > 1.0_dp) wdcoul = wdcoul + j%c%occ(i1)*(j%c%occ(i1) - 1.0_dp)*v(idx)/2.0_dp
       if (j%c%m(i1) > 0) idx = idx + 1
       do i2 = i1+1, j%c%norb
          idx = idx + 2
          k = j%c%k2(i1, i2)
          if (j%c%ilc(k) == 2) idx = idx + 1
       enddo
    enddo
    idx = 0
    do i1 = 1, j%c%norb
       idx = idx + 1
       if (j%c%m(i1) > 0) then
          idx = idx + 1
          if (abs(j%c%occ(i1) - 1.0_dp) > 1.e-7_dp) wex1 = wex1 + j%c%oxself(i1)*v(idx)
       endif
       do i2 = i1+1, j%c%norb
          k = j%c%k2(i1, i2)
          idx = idx + 1
          wndc = wndc + j%c%occ(i1)*j%c%occ(i2)*v(idx)
          idx = idx + 1
          wex1 = wex1 + j%c%ox1(i1,i2)*v(idx)
          if (j%c%ilc(k) == 2) then
             idx = idx + 1
             wex2 = wex2 + j%c%ox2(i1,i2)*v(idx)
          endif
       enddo
    enddo
    evt = woneel + wdcoul + wndc - wex1 - wex2
    epott = j%vnt + wdcoul + wndc - wex1 - wex2
    j%virial = (epott + j%c%z1*j%c%z2/j%c%r)/j%vkt
    j%etot = evt + j%c%z1*j%c%z2/j%c%r
    j%ecoul = wdcoul + wndc
    j%eexch = -wex1 - wex2
  end subroutine etotal_hf

  ! ---------------------------------------------------------------------------
  ! prepSCF for a start from guessed functions
  subroutine prep_scf(j)
    type(scf_job), intent(inout) :: j
    integer :: i, ii
    do i = 1, j%c%norb
       call fock_hf(j, i, j%f1o(i), j%f2o(i))
    enddo
    do ii = 1, j%c%norb
       i = j%c%norb + 1 - ii
       call norm(j, i)
       call ortho(j, i)
    enddo
    j%itouch = 1
    call mpole_moments(j, j%itouch)
    do i = j%c%norb, 1, -1
       call ea_hf(j, i)
       j%eeold(i) = j%ee(i,i)
    enddo
    call etotal_hf(j)
    call scf_init_state(j)
  end subroutine prep_scf

  ! prepSCF for a start from a checkpoint that carries the Lagrange
  ! multipliers (readFuncs leaves initAddData false, doSCF.F90:47-174): the
  ! Fock operators from the loaded orbitals, potentials and multipliers ee,
  ! no normalisation or orthogonalisation, then the total energy.  The
  ! multipole moments are formed from the loaded orbitals.
  subroutine prep_scf_restart(j, ee)
    type(scf_job), intent(inout) :: j
    real(dp), intent(in) :: ee(:,:)
    integer :: i, n
    n = j%c%norb
    j%ee(1:n, 1:n) = ee(1:n, 1:n)
    do i = 1, n
       call fock_hf(j, i, j%f1o(i), j%f2o(i))
    enddo
    j%itouch = 1
    call mpole_moments(j, j%itouch)
    call etotal_hf(j)
    call scf_init_state(j)
  end subroutine prep_scf_restart

  ! scf's own initialisation (doSCF.F90:265-334)
  subroutine scf_init_state(j)
    type(scf_job), intent(inout) :: j
    integer :: i
    j%iscf = 0
    j%nresets = 0; j%nresets4no = 0; j%norbenergythld = 0; j%norbnormthld = 0
    j%eeconv = 10.0_dp
    j%itouch = 0
    j%deltaeemax = 100.0_dp
    j%deltaee = 0.0_dp
    do i = 1, j%c%norb
       j%eeold(i) = j%ee(i,i)
    enddo
    j%deltanormmax = 100.0_dp
    j%lenergythld = .false.; j%lnormthld = .false.
    j%eeconvprev = 1.e35_dp; j%dnmaxprev = 1.e35_dp
    j%noenergydec = 0; j%nonormdec = 0
    j%done = .false.; j%stopcode = 0
    j%ctrl_dirty = .true.
    call setci(j, 0)
  end subroutine scf_init_state

  ! one orbital of the plain schedule (doSCF.F90:368-533)
  subroutine orbital_step(j, i)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    j%itouch(i) = 1
    if (j%iasympt /= 0 .and. j%ifixorb(i) == 0) then
       call pot_asympt(j, i)
       j%iasympt = j%iasympt - 1
    endif
    call relax_pots(j, i)
    if (.not. j%lfixorb) call relax_orb(j, i)
    if (j%ifixorb(i) == 0) call norm(j, i)
    call setci(j, i)
    if (j%c%norb > 1 .and. j%ifixorb(i) == 0) call ortho(j, i)
    if (j%ifixorb(i) /= 0) call fock_hf(j, i, j%f1o(i), j%f2o(i))
    call ea_hf(j, i)
    call energy_change(j, i)
  end subroutine orbital_step

  subroutine energy_change(j, i)
    type(scf_job), intent(inout) :: j
    integer, intent(in) :: i
    real(dp) :: thld
    thld = 10.0_dp**dble(-j%o%ienterm)
    j%deltaee(i) = j%ee(i,i) - j%eeold(i)
    if (abs(j%deltaee(i)) < eps0) j%deltaee(i) = 2.0_dp*thld
    j%eeold(i) = j%ee(i,i)
  end subroutine energy_change

  ! one SCF iteration of the plain schedule; returns .true. when the SCF stops
  logical function scf_iteration(j) result(stop_now)
    type(scf_job), intent(inout) :: j
    integer :: ii
    j%iscf = j%iscf + 1
    j%iasympt = j%c%norb
    do ii = 1, j%c%norb
       call orbital_step(j, j%c%norb - ii + 1)
    enddo
    stop_now = scf_control(j)
    j%ctrl_dirty = .true.
    if (j%mm_due) then
       call mpole_moments(j, j%mm_touch)
       j%mm_due = .false.
    endif
  end function scf_iteration

  ! The end-of-iteration decision: the reference's rules (convergence_control)
  ! plus the FP64 finish of a mixed-precision job.  A job relaxing with FP32
  ! correction sweeps converges to the fixed point of that iteration, which
  ! lies within rounding of the FP64 sweeps' fixed point but not on it (the
  ! FP64 update's own rounding shifts its fixed point by a few units in the
  ! last place, 1e-10 Eh in the total energy of FH on 481x901).  With polish
  ! > 0 the job switches to FP64 sweeps when it would otherwise stop (or
  ! `polish` iterations before maxscf) and runs at least `polish` more
  ! iterations under the same rules, so it ends where the FP64 path ends.
  logical function scf_control(j) result(stop_now)
    type(scf_job), intent(inout) :: j
    logical :: threshold_stop
    stop_now = convergence_control(j)
    if (j%o%polish <= 0) return
    threshold_stop = stop_now .and. j%stopcode /= 3 .and. j%stopcode < 6
    if (j%o%prec == 1) then
       if (threshold_stop .or. (.not. stop_now .and. j%iscf >= j%o%maxscf - j%o%polish)) then
          j%o%prec = 0
          j%polish_end = min(j%iscf + j%o%polish, j%o%maxscf)
          ! the reference's rules start over for the FP64 iterations
          j%lenergythld = .false.; j%lnormthld = .false.
          j%norbenergythld = 0; j%norbnormthld = 0; j%nresets = 0; j%nresets4no = 0
          j%noenergydec = 0; j%nonormdec = 0
          j%eeconvprev = 1.e35_dp; j%dnmaxprev = 1.e35_dp
          j%stopcode = 0
          stop_now = .false.
       endif
    else if (threshold_stop .and. j%iscf < j%polish_end) then
       j%stopcode = 0
       stop_now = .false.
    endif
  end function scf_control

  ! the end-of-iteration bookkeeping of scf (doSCF.F90:546-744)
  logical function convergence_control(j) result(stop_now)
    type(scf_job), intent(inout) :: j
    real(dp) :: ethld, nthld, thren4init
    integer :: nei
    stop_now = .false.
    ethld = 10.0_dp**dble(-j%o%ienterm)
    nthld = 10.0_dp**dble(-j%o%inoterm)
    thren4init = 10.0_dp**dble(-j%o%ienterm4init)

    j%deltaeemax = 0.0_dp
    do nei = 1, j%c%norb
       if (abs(j%deltaee(nei)) >= abs(j%deltaeemax)) j%deltaeemax = j%deltaee(nei)
    enddo
    j%deltanormmax = 0.0_dp
    do nei = 1, j%c%norb
       if (j%ifixorb(nei) == 0) then
          if (abs(j%orbnorm(nei) - 1.0_dp) >= j%deltanormmax) then
             j%deltanormmax = abs(j%orbnorm(nei) - 1.0_dp)
             if (abs(j%deltanormmax) < eps0) j%deltanormmax = eps0
          endif
       endif
    enddo
    if (abs(j%eeconv) < eps0) j%eeconv = j%deltaeemax

    ! failure guards.  A NaN anywhere in an orbital or in a potential it sees
    ! reaches that orbital's norm or energy, which the device quadratures
    ! form over every grid point (the reference can only report one,
    ! detectNaN detect.F90:9-25, relaxOrbs.F90:379-405); an orbital-energy
    ! change above orbEnergyIncThld is divergence (the test of
    ! doSCF.F90:612-616, applied to the largest change of this iteration).
    ! Either retires the job instead of iterating to maxscf.
    do nei = 1, j%c%norb
       if (ieee_is_nan(j%ee(nei,nei)) .or. ieee_is_nan(j%orbnorm(nei)) .or. ieee_is_nan(j%deltaee(nei))) then
          j%stopcode = 7; stop_now = .true.; return
       endif
    enddo
    if (abs(j%deltaeemax) > j%o%orb_energy_inc_thld) then
       j%stopcode = 6; stop_now = .true.; return
    endif

    ! saturation (no decrease for nenlast/nnolast iterations after nscf2skip)
    if (j%iscf > j%o%nscf2skip) then
       if (abs(j%deltaeemax) >= abs(j%eeconvprev)) then
          j%noenergydec = j%noenergydec + 1
       else
          j%eeconvprev = j%deltaeemax
          j%noenergydec = 0
       endif
       if (abs(j%deltanormmax) >= abs(j%dnmaxprev)) then
          j%nonormdec = j%nonormdec + 1
       else
          j%dnmaxprev = j%deltanormmax
          j%nonormdec = 0
       endif
       if (j%noenergydec > j%o%nenlast .or. j%nonormdec > j%o%nnolast) then
          j%stopcode = 4; stop_now = .true.; return
       endif
    else
       j%eeconvprev = 100.0_dp
       j%dnmaxprev = j%deltanormmax
    endif
    if (j%iscf == j%o%maxscf) then
       j%stopcode = 3; stop_now = .true.; return
    endif

    if (j%lfixorb4init .and. abs(j%deltaeemax) < thren4init) then
       if (j%norbenergythld == j%o%nscfextra) then
          j%ifixorb = 0
          j%lfixorb = .false.
          j%lfixorb4init = .false.
          j%eeconv = 10.0_dp
          j%deltaeemax = 100.0_dp
       else
          j%norbenergythld = j%norbenergythld + 1
       endif
    endif

    if (.not. j%lenergythld) then
       if (abs(j%deltaeemax) < ethld) then
          j%lenergythld = .true.
          j%eeconv = j%deltaeemax
          j%iasympt = j%c%norb
          j%itouch = 0
          j%norbenergythld = 0
       endif
    else
       if (abs(j%deltaeemax) < ethld) then
          if (j%norbenergythld == j%o%nscfextra) then
             j%stopcode = 1; stop_now = .true.; return
          else
             j%norbenergythld = j%norbenergythld + 1
          endif
       elseif (abs(j%deltanormmax) < nthld) then
          if (j%norbnormthld == j%o%nscfextra) then
             j%stopcode = 2; stop_now = .true.; return
          else
             j%norbnormthld = j%norbnormthld + 1
          endif
       else
          j%lenergythld = .false.
          if (j%nresets > j%o%maxresets) then
             j%stopcode = 5; stop_now = .true.; return
          else
             j%nresets = j%nresets + 1
          endif
       endif
    endif

    if (.not. j%lfixorb) then
       if (.not. j%lnormthld) then
          if (abs(j%deltanormmax) < nthld) then
             j%lnormthld = .true.
             j%iasympt = j%c%norb
             j%itouch = 0
             j%norbnormthld = 0
          endif
       else
          if (abs(j%deltanormmax) < nthld) then
             if (j%norbnormthld == j%o%nscfextra) then
                j%stopcode = 2; stop_now = .true.; return
             else
                j%norbnormthld = j%norbnormthld + 1
             endif
          else
             j%lnormthld = .false.
             if (j%nresets4no > j%o%maxresets) then
                j%stopcode = 5; stop_now = .true.; return
             else
                j%nresets4no = j%nresets4no + 1
             endif
          endif
       endif
    endif

    ! refresh the multipole moments when the energy change dropped enough
    if (abs(j%eeconv/j%deltaeemax) > j%o%recalc_mm) then
       j%eeconv = j%deltaeemax
       j%mm_due = .true.
       j%mm_touch = j%itouch
    endif
    j%iasympt = j%c%norb
    j%itouch = 0
  end function convergence_control

  ! the final pass of scf (doSCF.F90:793-858)
  subroutine scf_finish(j)
    type(scf_job), intent(inout) :: j
    real(dp) :: saved(MAXO)
    integer :: i, ii
    if (j%c%norb > 1) then
       saved = j%orbnorm
       do ii = 1, j%c%norb
          i = j%c%norb + 1 - ii
          call norm(j, i)
          call ortho(j, i)
       enddo
       j%orbnorm = saved
    endif
    do i = j%c%norb, 1, -1
       call ea_hf(j, i)
    enddo
    call mpole_moments(j, j%itouch)
    call etotal_hf(j)
    j%done = .true.
  end subroutine scf_finish

  ! ---------------------------------------------------------------------------
  ! One iteration of the batched schedule over several jobs.  Per job it keeps
  ! the same equations and control as the plain schedule, but relaxes all
  ! potentials of all jobs in one fused launch sequence, then all orbitals of
  ! all jobs, before the per-orbital normalisation, orthogonalisation and
  ! orbital energies.  Returns per job whether it stopped.
  subroutine batched_iteration(js, idx, nj, stopped)
    type(scf_job), intent(inout), target :: js(:)
    integer, intent(in) :: idx(:), nj
    logical, intent(inout) :: stopped(:)
    type(x2g_relax_item), allocatable :: items(:)
    type(x2g_tail_item), allocatable :: tails(:)
    type(x2g_normortho_item), allocatable :: no(:)
    type(x2g_eahf_item), allocatable :: ea(:)
    type(c_ptr), allocatable :: grids(:), recs(:), cptr(:), rptr(:)
    integer(c_int), allocatable :: cnt(:)
    integer, allocatable :: recq(:)
    real(dp), allocatable :: rech(:)
    logical, allocatable :: done(:)
    type(lc_batch) :: lb
    type(x2g_scf_ctrl) :: cc
    integer :: ndc
    type(x2g_term) :: t1(8), tc(MAXO), tx(4*MAXO)
    type(c_ptr) :: ctx
    integer :: q, h, i, ii, n, nt, nmax, ntmax, ms1, ms2, maxnorb, mpmax, i2, i3, jor, nno, nea, nrec
    integer :: n1, ntc, ntx, base
    nmax = 0; ntmax = 0; maxnorb = 0; mpmax = 2
    allocate(grids(nj))
    do q = 1, nj
       h = idx(q)
       grids(q) = js(h)%gd
       if (stopped(q)) cycle
       nmax = nmax + sum(js(h)%c%npot(1:js(h)%c%norb)) + js(h)%c%norb
       ntmax = ntmax + js(h)%c%norb + 2*js(h)%c%nslots
       maxnorb = max(maxnorb, js(h)%c%norb)
       mpmax = max(mpmax, js(h)%o%mpole)
    enddo
    if (nmax == 0) return
    ctx = js(idx(1))%ctx
    allocate(items(nmax), tails(max(1, ntmax)))

    ! 1. tails and right-hand sides of the potentials of every orbital of every
    ! job, then all potentials relaxed together
    n = 0; nt = 0
    do q = 1, nj
       if (stopped(q)) cycle
       h = idx(q)
       if (js(h)%o%host_control == 0 .and. js(h)%ctrl_dirty) call ctrl_upload(js(h))
       js(h)%iscf = js(h)%iscf + 1
       js(h)%iasympt = js(h)%c%norb
       allocate(done(max(1, js(h)%c%nslots)))
       done = .false.
       do ii = 1, js(h)%c%norb
          i = js(h)%c%norb - ii + 1
          js(h)%itouch(i) = 1
          if (js(h)%ifixorb(i) == 0) then
             call pot_tail_items(js(h), i, tails, nt, q-1, done)
             js(h)%iasympt = js(h)%iasympt - 1
          endif
          call pot_items(js(h), i, items, n, q-1, lb)
       enddo
       deallocate(done)
    enddo
    if (nt > 0) call x2check(x2g_pot_tail_batch(ctx, grids, tails, int(nt, c_int), int(mpmax, c_int)), &
         'batched potAsympt')
    call lc_run(lb, ctx, grids)
    ms1 = 0; ms2 = 0
    do q = 1, nj
       if (stopped(q)) cycle
       h = idx(q)
       ms1 = max(ms1, js(h)%o%maxsorpot(1)); ms2 = max(ms2, js(h)%o%maxsorpot(2))
    enddo
    call x2check(x2g_relax_batch(ctx, grids, items, int(n, c_int), int(ms1, c_int), int(ms2, c_int)), &
         'batched potentials')

    ! 2. Fock operators (Coulomb and exchange sums, then fock1, fock2 and the
    ! relaxation lhs = fock1 + diag), then all orbitals relaxed together
    do q = 1, nj
       h = idx(q)
       if (stopped(q) .or. js(h)%lfixorb) cycle
       if (js(h)%c%nel == 1) cycle
       do ii = 1, js(h)%c%norb
          i = js(h)%c%norb - ii + 1
          if (js(h)%ifixorb(i) == 1) cycle
          call fock_terms(js(h), i, tc, ntc, tx, ntx, t1, n1)
          call lc_add(lb, js(h)%cpot(i), c_null_ptr, 0.0_dp, tc, ntc, q-1)
          call lc_add(lb, js(h)%xpot(i), c_null_ptr, 0.0_dp, tx, ntx, q-1)
       enddo
    enddo
    call lc_run(lb, ctx, grids)
    n = 0
    ms1 = 0; ms2 = 0
    do q = 1, nj
       h = idx(q)
       if (stopped(q) .or. js(h)%lfixorb) cycle
       do ii = 1, js(h)%c%norb
          i = js(h)%c%norb - ii + 1
          if (js(h)%ifixorb(i) == 1) cycle
          call fock_terms(js(h), i, tc, ntc, tx, ntx, t1, n1)
          if (js(h)%c%nel == 1) then
             call lc_add(lb, js(h)%f1o(i), js(h)%S(5), 1.0_dp, t1, n1, q-1)
             call lc_add(lb, js(h)%f2o(i), c_null_ptr, 0.0_dp, t1, 0, q-1)
          else
             n1 = n1 + 1; t1(n1) = x2g_term(1.0_dp, js(h)%S(7), js(h)%cpot(i), c_null_ptr)
             call lc_add(lb, js(h)%f1o(i), js(h)%S(5), 1.0_dp, t1, n1, q-1)
             ! fock2 = S7 xpot: with the FP32 correction the residual forms it
             ! from the two factors (below)
             if (js(h)%o%prec /= 1) then
                tc(1) = x2g_term(1.0_dp, js(h)%S(7), js(h)%xpot(i), c_null_ptr)
                call lc_add(lb, js(h)%f2o(i), c_null_ptr, 0.0_dp, tc, 1, q-1)
             endif
          endif
          n1 = n1 + 1; t1(n1) = x2g_term(js(h)%g%diag, js(h)%ones, c_null_ptr, c_null_ptr)
          call lc_add(lb, js(h)%lo(i), js(h)%S(5), 1.0_dp, t1, n1, q-1)
          n = n + 1
          items(n) = x2g_relax_item(js(h)%orb(i), js(h)%lo(i), js(h)%f2o(i), js(h)%f1o(i), js(h)%omega_orb, &
               int(js(h)%c%isym(i), c_int), 0_c_int, int(js(h)%o%prec, c_int), int(q-1, c_int))
          call set_scheme(js(h), items(n), js(h)%o%orb_method, i)
          if (js(h)%o%prec == 1 .and. js(h)%c%nel /= 1) then
             items(n)%rhs_f1_d = js(h)%S(7)
             items(n)%rhs_f2_d = js(h)%xpot(i)
          endif
          js(h)%nsweep_orb = js(h)%nsweep_orb + int(js(h)%o%maxsororb(1)*js(h)%o%maxsororb(2), c_long_long)
       enddo
       ms1 = max(ms1, js(h)%o%maxsororb(1)); ms2 = max(ms2, js(h)%o%maxsororb(2))
    enddo
    call lc_run(lb, ctx, grids)
    if (n > 0) call x2check(x2g_relax_batch(ctx, grids, items, int(n, c_int), int(ms1, c_int), int(ms2, c_int)), &
         'batched orbitals')

    ! 3. normalisation, symmetry and orthogonalisation, orbital position by
    ! orbital position (each orbital against the higher ones of its symmetry)
    allocate(no(nj), ea(nj*maxnorb))
    do ii = 1, maxnorb
       nno = 0
       do q = 1, nj
          h = idx(q)
          if (stopped(q) .or. js(h)%c%norb < ii) cycle
          associate (j => js(h))
            i = j%c%norb - ii + 1
            if (j%ifixorb(i) /= 0 .and. .not. j%c%homonuclear) cycle
            nno = nno + 1
            no(nno) = x2g_normortho_item()
            no(nno)%psi_d = j%orb(i)
            no(nno)%normalise = merge(1_c_int, 0_c_int, j%ifixorb(i) == 0)
            no(nno)%ihomo = 0
            if (j%c%homonuclear) no(nno)%ihomo = int(j%c%ihomo(i), c_int)
            no(nno)%grid_index = int(q-1, c_int)
            no(nno)%norm_out_d = ptr_off(j%rec, 8*(i-1))
            jor = 0
            if (j%c%norb > 1 .and. j%ifixorb(i) == 0) then
               do i2 = 1, j%c%norb
                  i3 = j%c%norb - i2 + 1
                  if (i3 <= i) cycle
                  if (j%c%m(i3) /= j%c%m(i)) cycle
                  if (j%c%ige(i3) /= j%c%ige(i)) cycle
                  jor = jor + 1
                  no(nno)%others_d(jor) = j%orb(i3)
               enddo
            endif
            no(nno)%nothers = int(jor, c_int)
          end associate
       enddo
       if (nno > 0) call x2check(x2g_normortho_batch(ctx, grids, no, int(nno, c_int)), 'batched normOrtho')
       ! frozen orbitals refresh their Fock operators at this point of the order
       do q = 1, nj
          h = idx(q)
          if (stopped(q) .or. js(h)%c%norb < ii) cycle
          i = js(h)%c%norb - ii + 1
          if (js(h)%ifixorb(i) /= 0) call fock_hf(js(h), i, js(h)%f1o(i), js(h)%f2o(i))
       enddo
    enddo

    ! 4. orbital energies of every orbital of every job
    nea = 0
    do q = 1, nj
       h = idx(q)
       if (stopped(q)) cycle
       do i = 1, js(h)%c%norb
          nea = nea + 1
          ea(nea) = x2g_eahf_item(js(h)%orb(i), js(h)%cpot(i), js(h)%xpot(i), int(js(h)%c%isym(i), c_int), &
               int(js(h)%c%m(i), c_int), int(q-1, c_int), ptr_off(js(h)%rec, 8*(MAXO+i-1)))
       enddo
    enddo
    call x2check(x2g_eahf_batch(ctx, grids, ea, int(nea, c_int)), 'batched EaHF')

    ! 5. the end-of-iteration control of every job in one launch
    ! (x2g_scf_control, on the records the norms and energies went to), then
    ! one transfer of the control records; a job with host_control set
    ! transfers its norms and energies instead and runs the host rules
    allocate(recs(nj), cnt(nj), recq(nj), cptr(nj), rptr(nj))
    nrec = 0; ndc = 0
    do q = 1, nj
       if (stopped(q)) cycle
       h = idx(q)
       nrec = nrec + 1
       recq(nrec) = q
       if (js(h)%o%host_control == 0) then
          ndc = ndc + 1
          cptr(ndc) = js(h)%ctrl
          rptr(ndc) = js(h)%rec
          recs(nrec) = js(h)%ctrl
          cnt(nrec) = int(c_sizeof(cc)/8, c_int)
       else
          recs(nrec) = js(h)%rec
          cnt(nrec) = int(2*MAXO, c_int)
       endif
    enddo
    if (ndc > 0) call x2check(x2g_scf_control(ctx, cptr, rptr, int(ndc, c_int)), 'SCF control')
    allocate(rech(max(1, sum(cnt(1:nrec)))))
    call x2check(x2g_gather(ctx, recs, cnt, int(nrec, c_int), rech), 'gather')
    base = 0
    do n = 1, nrec
       q = recq(n)
       h = idx(q)
       if (js(h)%o%host_control == 0) then
          cc = transfer(rech(base + 1:base + cnt(n)), cc)
          call ctrl_unpack(js(h), cc)
          stopped(q) = cc%stop_flag /= 0
       else
          do ii = 1, js(h)%c%norb
             i = js(h)%c%norb - ii + 1
             if (js(h)%ifixorb(i) == 0) js(h)%orbnorm(i) = rech(base + i)
             js(h)%ee(i,i) = rech(base + MAXO + i)
             call energy_change(js(h), i)
          enddo
          stopped(q) = scf_control(js(h))
          js(h)%ctrl_dirty = .true.
       endif
       base = base + cnt(n)
    enddo
    call mm_refresh(js, idx, nj)
    deallocate(items, tails, grids, no, ea, recs, cnt, recq, rech, cptr, rptr)
  end subroutine batched_iteration

  ! The job's control state as the device control record (x2g_scf_ctrl):
  ! the settings of the rules and the state the host rules keep.
  subroutine ctrl_pack(j, c)
    type(scf_job), intent(in) :: j
    type(x2g_scf_ctrl), intent(out) :: c
    integer :: i
    c%ethld = 10.0_dp**dble(-j%o%ienterm)
    c%nthld = 10.0_dp**dble(-j%o%inoterm)
    c%thren4init = 10.0_dp**dble(-j%o%ienterm4init)
    c%recalc_mm = j%o%recalc_mm
    c%inc_thld = j%o%orb_energy_inc_thld
    do i = 1, MAXO
       c%ee(i) = j%ee(i,i)
    enddo
    c%eeold = j%eeold; c%deltaee = j%deltaee; c%orbnorm = j%orbnorm
    c%eeconv = j%eeconv; c%eeconvprev = j%eeconvprev; c%dnmaxprev = j%dnmaxprev
    c%deltaeemax = j%deltaeemax; c%deltanormmax = j%deltanormmax
    c%norb = j%c%norb; c%maxscf = j%o%maxscf; c%nscf2skip = j%o%nscf2skip
    c%nenlast = j%o%nenlast; c%nnolast = j%o%nnolast; c%nscfextra = j%o%nscfextra
    c%maxresets = j%o%maxresets; c%polish = j%o%polish
    c%ifixorb = j%ifixorb; c%itouch = j%itouch; c%mm_touch = j%mm_touch
    c%iscf = j%iscf; c%nresets = j%nresets; c%nresets4no = j%nresets4no
    c%norbenergythld = j%norbenergythld; c%norbnormthld = j%norbnormthld
    c%noenergydec = j%noenergydec; c%nonormdec = j%nonormdec
    c%lenergythld = merge(1, 0, j%lenergythld); c%lnormthld = merge(1, 0, j%lnormthld)
    c%lfixorb = merge(1, 0, j%lfixorb); c%lfixorb4init = merge(1, 0, j%lfixorb4init)
    c%mm_due = merge(1, 0, j%mm_due); c%prec = j%o%prec; c%polish_end = j%polish_end
    c%stopcode = j%stopcode; c%stop_flag = 0
  end subroutine ctrl_pack

  ! the state of a device control record back into the job
  subroutine ctrl_unpack(j, c)
    type(scf_job), intent(inout) :: j
    type(x2g_scf_ctrl), intent(in) :: c
    integer :: i
    do i = 1, MAXO
       j%ee(i,i) = c%ee(i)
    enddo
    j%eeold = c%eeold; j%deltaee = c%deltaee; j%orbnorm = c%orbnorm
    j%eeconv = c%eeconv; j%eeconvprev = c%eeconvprev; j%dnmaxprev = c%dnmaxprev
    j%deltaeemax = c%deltaeemax; j%deltanormmax = c%deltanormmax
    j%ifixorb = c%ifixorb; j%itouch = c%itouch; j%mm_touch = c%mm_touch
    j%iscf = c%iscf; j%nresets = c%nresets; j%nresets4no = c%nresets4no
    j%norbenergythld = c%norbenergythld; j%norbnormthld = c%norbnormthld
    j%noenergydec = c%noenergydec; j%nonormdec = c%nonormdec
    j%lenergythld = c%lenergythld /= 0; j%lnormthld = c%lnormthld /= 0
    j%lfixorb = c%lfixorb /= 0; j%lfixorb4init = c%lfixorb4init /= 0
    j%mm_due = c%mm_due /= 0; j%o%prec = c%prec; j%polish_end = c%polish_end
    j%stopcode = c%stopcode
    j%iasympt = j%c%norb
  end subroutine ctrl_unpack

  ! write the job's device control record from its host state
  subroutine ctrl_upload(j)
    type(scf_job), intent(inout) :: j
    type(x2g_scf_ctrl), target :: c
    call ctrl_pack(j, c)
    call x2check(x2g_h2d(j%ctx, j%ctrl, c_loc(c), c_sizeof(c)), 'x2g_h2d')
    j%ctrl_dirty = .false.
  end subroutine ctrl_upload

  ! The host bookkeeping of the SCF loop (energy_change + convergence_control)
  ! driven by given per-iteration orbital energies ee(norb, nit) and norms
  ! norms(norb, nit) instead of device results; no device call.  Returns the
  ! stop code (0: still running after nit iterations) and the iteration it
  ! stopped at.  Used to test the stop rules and the failure guards.
  subroutine control_probe(norb, nit, ee, norms, ee0, o, stopcode, iters)
    integer, intent(in) :: norb, nit
    real(dp), intent(in) :: ee(norb, nit), norms(norb, nit), ee0(norb)
    type(scf_opts), intent(in) :: o
    integer, intent(out) :: stopcode, iters
    type(scf_job) :: j
    integer :: it, i
    j%o = o
    j%c%norb = norb
    j%iscf = 0
    do i = 1, norb
       j%ee(i,i) = ee0(i)
       j%eeold(i) = ee0(i)
    enddo
    stopcode = 0
    iters = 0
    do it = 1, nit
       j%iscf = it
       do i = norb, 1, -1
          j%orbnorm(i) = norms(i, it)
          j%ee(i,i) = ee(i, it)
          call energy_change(j, i)
       enddo
       iters = it
       if (scf_control(j)) then
          stopcode = j%stopcode
          return
       endif
    enddo
  end subroutine control_probe
end module x2f_scf
