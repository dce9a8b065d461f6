! C-callable entry points of the Fortran SCF driver (libx2scf.so).  Jobs live
! in a fixed table so that several host threads may step different jobs at
! the same time; the table itself is only modified by x2f_job_create and
! x2f_job_destroy, which the caller serialises.
module x2f_api
  use, intrinsic :: iso_c_binding
  use x2f_capi
  use x2f_grid
  use x2f_case
  use x2f_scf
  implicit none
  private

  integer, parameter :: dp = c_double
  integer, parameter :: MAXJOBS = 256

  ! job specification (mirrored by x2dhf_amd/scf.py: JobSpec)
  type, bind(c) :: x2f_spec
     integer(c_int) :: nni, nmu, norb, maxscf, ienterm, inoterm, nscfextra, ienterm4init
     integer(c_int) :: maxsororb(2), maxsorpot(2), mpole, prec, schedule, device
     real(c_double) :: z1, z2, r, rinf, omega_orb, omega_pot, recalc_mm
     integer(c_int) :: m(MAXO), gu(MAXO), n1(MAXO), l1(MAXO), n2(MAXO), l2(MAXO)
     real(c_double) :: occ(MAXO), c1(MAXO), c2(MAXO), za(MAXO), zb(MAXO)
     ! saturation test of the conv card (inputInterface.F90:1079-1102); <= 0: default
     integer(c_int) :: nscf2skip, nenlast, nnolast
     ! FP64 finish of a mixed-precision job (scf_opts%polish); 0 = none
     integer(c_int) :: polish
     ! relaxation scheme of the FP64 sweeps: orbitals / potentials 0 MCSOR,
     ! 1 SOR; mesh ordering: an X2G_ORDER_* code, 0 col-wise .. 9 middle2 (x2g_relax_item)
     integer(c_int) :: orb_method, pot_method, ordering
     ! altsweeps (doSCF.F90:403-411): 1 = SOR sweeps backward for odd orbitals
     integer(c_int) :: alt_sweeps
     ! 1: the batched schedule runs its end-of-iteration control on the host
     ! (the default 0 runs it on the device, x2g_scf_control)
     integer(c_int) :: host_control
  end type x2f_spec

  ! job results (mirrored by x2dhf_amd/scf.py: JobResult)
  type, bind(c) :: x2f_result
     integer(c_int) :: status, iterations, stopcode, norb
     integer(c_long_long) :: nsweep_orb, nsweep_pot, bytes
     real(c_double) :: etot, ekin, enuc, ecoul, eexch, virial, delta_e, delta_norm
     real(c_double) :: eps(MAXO), norm(MAXO), cmulti(8, MAXO)
  end type x2f_result

  type(scf_job), allocatable, target, save :: jobs(:)
  logical, save :: used(MAXJOBS) = .false.

contains

  ! Build the grid coefficient arrays on the host (tests, benchmarks).
  integer(c_int) function x2f_grid_build(nni, nmu, z1, z2, r, rinf, coef, suppl, vec) &
       bind(c, name='x2f_grid_build') result(rc)
    integer(c_int), value :: nni, nmu
    real(c_double), value :: z1, z2, r, rinf
    real(c_double), intent(out) :: coef(25), suppl(*), vec(*)
    type(grid_t) :: g
    integer :: n
    rc = int(grid_build(g, int(nni), int(nmu), z1, z2, r, rinf), c_int)
    if (rc /= 0) return
    n = nni*nmu
    call grid_coef(g, coef)
    suppl(1:14*n) = reshape(g%suppl, [14*n])
    call grid_vec(g, vec(1:2*nmu+2*nni))
  end function x2f_grid_build

  subroutine x2f_omega_defaults(nni, nmu, z1, z2, oorb, opot) bind(c, name='x2f_omega_defaults')
    integer(c_int), value :: nni, nmu
    real(c_double), value :: z1, z2
    real(c_double), intent(out) :: oorb, opot
    call omega_defaults(int(nni), int(nmu), z1, z2, oorb, opot)
  end subroutine x2f_omega_defaults

  ! Create a job.  share > 0 makes the job use the device context of job
  ! `share` (required for jobs stepped together by x2f_jobs_iterate_batched).
  ! Errors: 1 bad grid, 2 unsupported case, 3 job table full, 4 bad share.
  integer(c_int) function x2f_job_create(spec, share, handle) bind(c, name='x2f_job_create') result(rc)
    type(x2f_spec), intent(in) :: spec
    integer(c_int), value :: share
    integer(c_int), intent(out) :: handle
    type(c_ptr) :: ctx_in
    handle = -1
    ctx_in = c_null_ptr
    if (.not. allocated(jobs)) allocate(jobs(MAXJOBS))
    if (share > 0) then
       if (.not. valid(share)) then
          rc = 4; return
       endif
       ctx_in = jobs(share)%ctx
    endif
    rc = create_job(spec, ctx_in, handle)
  end function x2f_job_create

  ! Create a job on an x2g_ctx owned by the caller (it must outlive the job);
  ! jobs on one context can be stepped together.  Errors as x2f_job_create.
  integer(c_int) function x2f_job_create_ctx(spec, ctx, handle) bind(c, name='x2f_job_create_ctx') result(rc)
    type(x2f_spec), intent(in) :: spec
    type(c_ptr), value :: ctx
    integer(c_int), intent(out) :: handle
    handle = -1
    if (.not. c_associated(ctx)) then
       rc = 4; return
    endif
    if (.not. allocated(jobs)) allocate(jobs(MAXJOBS))
    rc = create_job(spec, ctx, handle)
  end function x2f_job_create_ctx

  ! Device bytes a job of this specification holds (its arrays; the shared
  ! library workspace is not included), without creating it; -1 for a
  ! specification that x2f_job_create would reject.
  integer(c_long_long) function x2f_spec_bytes(spec) bind(c, name='x2f_spec_bytes')
    type(x2f_spec), intent(in) :: spec
    type(case_t) :: c
    integer :: i, ierr, npotall
    integer(c_long_long) :: n
    x2f_spec_bytes = -1
    if (mod(spec%nni - 1, 6) /= 0 .or. mod(spec%nmu - 1, 6) /= 0 .or. spec%nni < 13 .or. spec%nmu < 13) return
    c = case_t()
    c%norb = spec%norb
    c%z1 = spec%z1; c%z2 = spec%z2; c%r = spec%r
    do i = 1, min(int(spec%norb), MAXO)
       c%m(i) = spec%m(i); c%gu(i) = spec%gu(i); c%occ(i) = spec%occ(i)
    enddo
    ierr = case_setup(c)
    if (ierr /= 0) return
    npotall = sum(c%npot(1:c%norb))
    n = int(spec%nni, c_long_long)*int(spec%nmu, c_long_long)
    ! the arrays of job_init (orbitals, potentials, Fock parts, lhs/rhs of
    ! every potential, ones, record) plus the 14 grid coefficient arrays
    x2f_spec_bytes = 8_c_long_long*n*int(7*c%norb + c%nslots + 2*npotall + 2 + 14, c_long_long)
  end function x2f_spec_bytes

  ! Device workspace bytes a job of this specification adds to its context
  ! while it is relaxed (x2g_relax_workspace_bytes of its largest relaxation
  ! call, the potentials, plus the padded copies of the batched quadratures
  ! and their partial sums), without creating it; -1 as x2f_spec_bytes.
  integer(c_long_long) function x2f_spec_workspace_bytes(spec) bind(c, name='x2f_spec_workspace_bytes')
    type(x2f_spec), intent(in) :: spec
    type(case_t) :: c
    integer :: i, ierr, npotall
    integer(c_long_long) :: per, npad
    x2f_spec_workspace_bytes = -1
    if (x2f_spec_bytes(spec) < 0) return
    c = case_t()
    c%norb = spec%norb
    c%z1 = spec%z1; c%z2 = spec%z2; c%r = spec%r
    do i = 1, min(int(spec%norb), MAXO)
       c%m(i) = spec%m(i); c%gu(i) = spec%gu(i); c%occ(i) = spec%occ(i)
    enddo
    ierr = case_setup(c)
    if (ierr /= 0) return
    npotall = sum(c%npot(1:c%norb))
    per = x2g_relax_workspace_bytes(spec%nni, spec%nmu, spec%prec)
    if (per < 0) return
    npad = int(spec%nni + 8, c_long_long)*int(spec%nmu + 8, c_long_long)
    x2f_spec_workspace_bytes = int(max(npotall, int(spec%norb)), c_long_long)*per + &
         int(spec%norb, c_long_long)*(8_c_long_long*npad + 65536_c_long_long)
  end function x2f_spec_workspace_bytes

  ! The SCF loop's host bookkeeping (energy_change, convergence_control) fed
  ! with given orbital energies ee(norb, nit) and norms(norb, nit), starting
  ! from energies ee0(norb); no device call.  Returns the stop code (0 still
  ! running after nit iterations, else as x2f_result's) in stopcode and the
  ! iteration it stopped at in iters.
  integer(c_int) function x2f_control_probe(norb, nit, ee, norms, ee0, maxscf, ienterm, inoterm, nscf2skip, &
       stopcode, iters) bind(c, name='x2f_control_probe') result(rc)
    integer(c_int), value :: norb, nit, maxscf, ienterm, inoterm, nscf2skip
    real(c_double), intent(in) :: ee(norb, nit), norms(norb, nit), ee0(norb)
    integer(c_int), intent(out) :: stopcode, iters
    type(scf_opts) :: o
    integer :: sc, it
    rc = 1
    stopcode = 0; iters = 0
    if (norb < 1 .or. norb > MAXO .or. nit < 1) return
    o = scf_opts()
    o%maxscf = maxscf; o%ienterm = ienterm; o%inoterm = inoterm
    o%ienterm4init = 1
    if (nscf2skip > 0) o%nscf2skip = nscf2skip
    call control_probe(int(norb), int(nit), ee, norms, ee0, o, sc, it)
    stopcode = int(sc, c_int); iters = int(it, c_int)
    rc = 0
  end function x2f_control_probe

  integer(c_int) function create_job(spec, ctx_in, handle) result(rc)
    type(x2f_spec), intent(in) :: spec
    type(c_ptr), intent(in) :: ctx_in
    integer(c_int), intent(out) :: handle
    integer :: h, i, ierr
    real(dp) :: s
    h = 0
    do i = 1, MAXJOBS
       if (.not. used(i)) then
          h = i; exit
       endif
    enddo
    if (h == 0) then
       rc = 3; return
    endif
    associate (j => jobs(h))
      j%c = case_t()
      j%c%norb = spec%norb
      j%c%z1 = spec%z1; j%c%z2 = spec%z2; j%c%r = spec%r
      do i = 1, min(int(spec%norb), MAXO)
         j%c%m(i) = spec%m(i); j%c%gu(i) = spec%gu(i); j%c%occ(i) = spec%occ(i)
         ! mixing coefficients normalised to |c1|+|c2| = 1 (inputInterface.F90:1867-1874)
         s = abs(spec%c1(i)) + abs(spec%c2(i))
         if (s < epsilon(1.0_dp)) s = 1.0_dp
         j%c%c1(i) = spec%c1(i)/s; j%c%c2(i) = spec%c2(i)/s
         j%c%za(i) = spec%za(i); j%c%zb(i) = spec%zb(i)
         j%c%n1(i) = spec%n1(i); j%c%l1(i) = spec%l1(i); j%c%n2(i) = spec%n2(i); j%c%l2(i) = spec%l2(i)
      enddo
      ierr = case_setup(j%c)
      if (ierr /= 0) then
         rc = 2; return
      endif
      ierr = grid_build(j%g, int(spec%nni), int(spec%nmu), spec%z1, spec%z2, spec%r, spec%rinf)
      if (ierr /= 0) then
         rc = 1; return
      endif
      j%o = scf_opts()
      j%o%maxscf = spec%maxscf; j%o%ienterm = spec%ienterm; j%o%inoterm = spec%inoterm
      j%o%nscfextra = spec%nscfextra; j%o%ienterm4init = spec%ienterm4init
      j%o%maxsororb = spec%maxsororb; j%o%maxsorpot = spec%maxsorpot
      j%o%mpole = spec%mpole; j%o%prec = spec%prec; j%o%schedule = spec%schedule
      j%o%omega_orb = spec%omega_orb; j%o%omega_pot = spec%omega_pot; j%o%recalc_mm = spec%recalc_mm
      if (spec%polish > 0) j%o%polish = spec%polish
      j%o%orb_method = spec%orb_method; j%o%pot_method = spec%pot_method; j%o%ordering = spec%ordering
      j%o%alt_sweeps = spec%alt_sweeps
      j%o%host_control = spec%host_control
      if (spec%nscf2skip > 0) j%o%nscf2skip = spec%nscf2skip
      if (spec%nenlast > 0) j%o%nenlast = spec%nenlast
      if (spec%nnolast > 0) j%o%nnolast = spec%nnolast
      ierr = job_init(j, int(spec%device), ctx_in)
      used(h) = .true.
      handle = h
      rc = 0
    end associate
  end function create_job

  logical function valid(handle)
    integer(c_int), intent(in) :: handle
    valid = handle >= 1 .and. handle <= MAXJOBS .and. allocated(jobs)
    if (valid) valid = used(handle)
  end function valid

  ! Set-up pass (prepSCF).
  integer(c_int) function x2f_job_prep(handle) bind(c, name='x2f_job_prep') result(rc)
    integer(c_int), value :: handle
    rc = 1
    if (.not. valid(handle)) return
    call prep_scf(jobs(handle))
    rc = 0
  end function x2f_job_prep

  ! Set-up pass after a checkpoint was loaded with x2f_job_put: the Lagrange
  ! multipliers ee(ld, norb) come from the checkpoint, the orbitals are taken
  ! as they are (doSCF.F90:47-174 with initAddData false).
  integer(c_int) function x2f_job_prep_restart(handle, ee, ld) bind(c, name='x2f_job_prep_restart') result(rc)
    integer(c_int), value :: handle, ld
    real(c_double), intent(in) :: ee(ld, *)
    integer :: n
    rc = 1
    if (.not. valid(handle)) return
    n = jobs(handle)%c%norb
    if (ld < n) return
    call prep_scf_restart(jobs(handle), ee(1:n, 1:n))
    rc = 0
  end function x2f_job_prep_restart

  ! Up to niter iterations of the job's own schedule; 1 when the SCF stopped.
  integer(c_int) function x2f_job_iterate(handle, niter) bind(c, name='x2f_job_iterate') result(rc)
    integer(c_int), value :: handle, niter
    integer :: k
    logical :: st(1)
    rc = -1
    if (.not. valid(handle)) return
    rc = 0
    do k = 1, niter
       if (jobs(handle)%o%schedule == 1) then
          st(1) = .false.
          call batched_iteration(jobs, [int(handle)], 1, st)
          if (st(1)) then
             rc = 1; exit
          endif
       else
          if (scf_iteration(jobs(handle))) then
             rc = 1; exit
          endif
       endif
    enddo
  end function x2f_job_iterate

  ! One batched iteration over n jobs that share a device context; stopped(i)
  ! is set to 1 for the jobs whose SCF stopped.
  integer(c_int) function x2f_jobs_iterate_batched(n, handles, stopped) &
       bind(c, name='x2f_jobs_iterate_batched') result(rc)
    integer(c_int), value :: n
    integer(c_int), intent(in) :: handles(n)
    integer(c_int), intent(inout) :: stopped(n)
    logical, allocatable :: st(:)
    integer, allocatable :: idx(:)
    integer :: q
    rc = -1
    do q = 1, n
       if (.not. valid(handles(q))) return
       if (.not. c_associated(jobs(handles(q))%ctx, jobs(handles(1))%ctx)) return
    enddo
    allocate(idx(n), st(n))
    idx = handles
    st = stopped /= 0
    call batched_iteration(jobs, idx, int(n), st)
    stopped = merge(1_c_int, 0_c_int, st)
    deallocate(idx, st)
    rc = 0
  end function x2f_jobs_iterate_batched

  integer(c_int) function x2f_job_finish(handle) bind(c, name='x2f_job_finish') result(rc)
    integer(c_int), value :: handle
    rc = 1
    if (.not. valid(handle)) return
    call scf_finish(jobs(handle))
    rc = 0
  end function x2f_job_finish

  ! prep + iterations to the end + final pass
  integer(c_int) function x2f_job_run(handle) bind(c, name='x2f_job_run') result(rc)
    integer(c_int), value :: handle
    rc = x2f_job_prep(handle)
    if (rc /= 0) return
    do
       rc = x2f_job_iterate(handle, 1000000_c_int)
       if (rc /= 0) exit
    enddo
    rc = x2f_job_finish(handle)
  end function x2f_job_run

  integer(c_int) function x2f_job_result(handle, res) bind(c, name='x2f_job_result') result(rc)
    integer(c_int), value :: handle
    type(x2f_result), intent(out) :: res
    integer :: i
    rc = 1
    if (.not. valid(handle)) return
    associate (j => jobs(handle))
      res%status = merge(1_c_int, 0_c_int, j%done)
      res%iterations = j%iscf
      res%stopcode = j%stopcode
      res%norb = j%c%norb
      res%nsweep_orb = j%nsweep_orb
      res%nsweep_pot = j%nsweep_pot
      res%bytes = int(job_bytes(j), c_long_long)
      res%etot = j%etot; res%ekin = j%vkt; res%enuc = j%vnt
      res%ecoul = j%ecoul; res%eexch = j%eexch; res%virial = j%virial
      res%delta_e = j%deltaeemax; res%delta_norm = j%deltanormmax
      res%eps = 0; res%norm = 0; res%cmulti = 0
      do i = 1, j%c%norb
         res%eps(i) = j%ee(i,i)
         res%norm(i) = j%orbnorm(i)
         res%cmulti(:, i) = j%cmulti(:, i)
      enddo
    end associate
    rc = 0
  end function x2f_job_result

  ! Copy orbital i (kind 0), Coulomb potential i (kind 1) or exchange slot i
  ! (kind 2) to the host array out(nni*nmu).
  integer(c_int) function x2f_job_get(handle, kind, i, out) bind(c, name='x2f_job_get') result(rc)
    integer(c_int), value :: handle, kind, i
    real(c_double), intent(out), target :: out(*)
    type(c_ptr) :: p
    rc = 1
    if (.not. valid(handle)) return
    associate (j => jobs(handle))
      select case (kind)
      case (0)
         if (i < 1 .or. i > j%c%norb) return
         p = j%orb(i)
      case (1)
         if (i < 1 .or. i > j%c%norb) return
         p = j%coul(i)
      case (2)
         if (i < 1 .or. i > j%c%nslots) return
         p = j%exch(i)
      case default
         return
      end select
      rc = x2g_d2h(j%ctx, c_loc(out), p, 8_c_size_t*j%n)
    end associate
  end function x2f_job_get

  ! Copy the host array in(nni*nmu) to orbital i (kind 0), Coulomb potential i
  ! (kind 1) or exchange slot i (kind 2): loading a checkpoint (readFuncs,
  ! diskInterface.F90:247-779).  Call before x2f_job_prep, which forms the
  ! Fock operators, norms, moments and energies from what is loaded.
  integer(c_int) function x2f_job_put(handle, kind, i, in) bind(c, name='x2f_job_put') result(rc)
    integer(c_int), value :: handle, kind, i
    real(c_double), intent(in), target :: in(*)
    type(c_ptr) :: p
    rc = 1
    if (.not. valid(handle)) return
    associate (j => jobs(handle))
      select case (kind)
      case (0)
         if (i < 1 .or. i > j%c%norb) return
         p = j%orb(i)
      case (1)
         if (i < 1 .or. i > j%c%norb) return
         p = j%coul(i)
      case (2)
         if (i < 1 .or. i > j%c%nslots) return
         p = j%exch(i)
      case default
         return
      end select
      rc = x2g_h2d(j%ctx, p, c_loc(in), 8_c_size_t*j%n)
      if (rc == 0) rc = x2g_ctx_sync(j%ctx)
    end associate
  end function x2f_job_put

  ! Device context of a job (x2g_ctx*, for x2g_stats_* and x2g_ctx_sync);
  ! null for an invalid handle.
  type(c_ptr) function x2f_job_ctx(handle) bind(c, name='x2f_job_ctx')
    integer(c_int), value :: handle
    x2f_job_ctx = c_null_ptr
    if (.not. valid(handle)) return
    x2f_job_ctx = jobs(handle)%ctx
  end function x2f_job_ctx

  integer(c_long_long) function x2f_job_bytes(handle) bind(c, name='x2f_job_bytes')
    integer(c_int), value :: handle
    x2f_job_bytes = -1
    if (.not. valid(handle)) return
    x2f_job_bytes = int(job_bytes(jobs(handle)), c_long_long)
  end function x2f_job_bytes

  ! Destroying the job that created a shared context while other jobs still
  ! use it hands the context to one of them, which destroys it in turn; the
  ! context ends with the last job that uses it, in any order of destruction.
  integer(c_int) function x2f_job_destroy(handle) bind(c, name='x2f_job_destroy') result(rc)
    integer(c_int), value :: handle
    integer :: k
    rc = 1
    if (.not. valid(handle)) return
    if (jobs(handle)%own_ctx) then
       do k = 1, MAXJOBS
          if (k == handle .or. .not. used(k)) cycle
          if (c_associated(jobs(k)%ctx, jobs(handle)%ctx)) then
             jobs(k)%own_ctx = .true.
             jobs(handle)%own_ctx = .false.
             exit
          endif
       enddo
    endif
    call job_free(jobs(handle))
    used(handle) = .false.
    rc = 0
  end function x2f_job_destroy
end module x2f_api
