! Golden-vector harness (test infrastructure only).
!
! Drives the reference x2dhf routines, compiled from their own sources by
! oracle/ref.mk, through the set-up sequence of the reference main program
! (x2dhf.F90:77-199: setPrecision, setDefaults, inputData, arena allocation,
! initArrays, initOrbPot, prepSCF) and then through the first orbital step of
! one SCF iteration (doSCF.F90:368-533), dumping every array the relaxation
! path reads or writes as raw little-endian float64 / int32 streams into the
! directory named by the X2_DUMP environment variable.  tests/golden/make_golden.py
! turns those dumps into the small .npz fixtures the test-suite checks.
!
! The arena sizes below restate x2dhf.F90:95-132 (the reference main program is
! not linked: this program replaces it).
program ref_harness
  use params
  use commons
  use discrete
  use solver
  use scfshr
  use memory
  use sharedMemory
  use initCBAllocArrays
  use initOrbitalsPotentials
  use initVariables
  use inputInterface
  use doSCF
  use fock
  use inout
  use sormcsor
  use relaxOrbs
  use relaxPots
  use normOrtho
  use lagrangeMultipliers
  use integrals
  use totalEnergy
  use scfUtils
  use utils
  implicit none

  integer (KIND=IPREC) :: ni, mu, no, nons, ni8, mu8, ngrid, ngrid8
  integer (KIND=IPREC8) :: n
  integer (KIND=IPREC) :: iorb, i, j, k, ip, isym, in1, in2, idel, ibexp, nsw
  character(len=512) :: ddir
  character(len=16) :: siorb
  character(len=3) :: pfx
  real (PREC), allocatable :: lhs(:), rhs(:), wk(:), wk0(:), wk1(:), wk2(:), wk3(:)
  real (PREC) :: woneel, wtw, xv(8)
  integer :: iu

  call get_environment_variable('X2_DUMP', ddir)
  if (len_trim(ddir) == 0) ddir = 'dump'

  nsuppl = 14
  nsctch = 14
  call setPrecision
  call setDefaults
  call inputData(ni, mu, no, nons)
  firstOrb = 1
  lastOrb = norb

  ni8 = ni + 8
  mu8 = mu + 8
  ngrid = ni*mu
  ngrid8 = ni8*mu8
  length1 = no*ngrid
  length2 = (no+2)*ngrid
  length3 = no*ngrid + (no*(no+1)/2 + nons*(nons+1)/2)*ngrid8
  if (lcoulexch) length3 = length3 + length2
  length4 = nsuppl*ngrid
  length5 = nsctch*ngrid8
  length6 = 2*ngrid8 + 4*ni*(mu+16) + 2*maxgrids*(ni+mu)
  length7 = 30*ngrid8
  length8 = 8*ngrid

  allocate(orbptr(length1), exchptr(length3), supplptr(length4), scratchptr(length5))
  allocate(sorptr(length6), scratch4lxcptr(length7), coulombptr(length1))
  allocate(exchangeptr(length1), legendreptr(length8))
  orbptr = 0; exchptr = 0; supplptr = 0; scratchptr = 0; scratch4lxcptr = 0
  coulombptr = 0; exchangeptr = 0; legendreptr = 0
  do n = 1, length6
     sorptr(n) = 0
  enddo
  ee = 0

  call initArrays
  call initOrbPot

  ! ---- grid and coefficient arrays (initCBlocks / initSuppl)
  call dumpd('vmu', mxnmu, vmu)
  call dumpd('vxi', mxnmu, vxi)
  call dumpd('vni', nni, vni)
  call dumpd('veta', nni, veta)
  call dumpd('suppl', 14*mxsize, supplptr)
  call dumpd('legendre', 8*mxsize, legendreptr)
  xv(1) = hni; xv(2) = hmu(1); xv(3) = diag; xv(4) = ovforb; xv(5) = ovfcoul; xv(6) = ovfexch
  xv(7) = r; xv(8) = rinf
  call dumpd('scal', 8, xv)
  call dumpd('dmu1', 4, dmu1); call dumpd('dmu2', 4, dmu2)
  call dumpd('dni1', 4, dni1); call dumpd('dni2', 4, dni2)
  call dumpd('exeven', 5, exeven)
  call dumpi('sizes', 4, (/ nni, mxnmu, norb, int(length3) /))

  ! ---- initial state (initOrbPot)
  call dumpd('orb_init', norb*mxsize, orbptr)
  call dumpd('exch_init', int(length3), exchptr)

  call prepSCF
  call dumpd('orb_prep', norb*mxsize, orbptr)
  call dumpd('exch_prep', int(length3), exchptr)
  call dumpd('ee_prep', maxorb*maxorb, ee)
  call dumpd('cmulti_prep', maxorb*maxmpole, cmulti)
  call dumpd('excdi_prep', maxorb*(maxorb+1), excdi)
  call dumpd('excqu_prep', maxorb*(maxorb+1), excqu)
  call dumpd('excoc_prep', maxorb*(maxorb+1), excoc)
  call dumpd('exche_prep', maxorb*(maxorb+1), exche)
  xv = 0; xv(1) = etot; xv(2) = vkt; xv(3) = vnt; xv(4) = evt; xv(5) = encoul; xv(6) = enexch
  call dumpd('etot_prep', 8, xv)
  call dumpd('norm_prep', maxorb, orbNorm)

  ! ---- one orbital step of the first SCF iteration: iorb = norb
  if (lhomonucl .and. .not. lbreakCi) call setci(izero, orbptr)
  ! the orbital whose step is traced: X2_IORB (default norb, the first one
  ! the reference relaxes in an SCF iteration)
  iorb = norb
  call get_environment_variable('X2_IORB', siorb)
  if (len_trim(siorb) > 0) read(siorb, *) iorb
  do i = 1, norb
     itouch(i) = 0
  enddo
  itouch(iorb) = 1
  isstart = 1; isstop = ngrid1; isstep = 1
  ! a hydrogenic start keeps the orbitals frozen until the potentials settle
  ! (ifixorb/lfixorb4init); the traced step relaxes its orbital regardless
  ifixorb = 0
  lfixorb = .false.

  ! asymptotic (multipole) boundary values of the potentials of iorb
  call potAsympt(iorb, exchptr, exchptr)
  call dumpd('exch_asympt', int(length3), exchptr)
  call dumpi('potlist', 5*maxorb*2, (/ (ins1(iorb,k), ins2(iorb,k), deltam4pot(iorb,k), &
       isyms(iorb,k), ibexcp(iorb,k), k=1,2*maxorb) /))
  call dumpi('nexchpots', 1, (/ nexchpots(iorb) /))

  ! raw kernel goldens: the first two potential problems of iorb, one mcsor
  ! call of maxsor2 sweeps, one of a single sweep, one sor call and one
  ! backward sor call each, on identical input (prefix k1_ / k2_)
  allocate(lhs(mxsize), rhs(mxsize), wk(mxsize8), wk0(mxsize8))
  do ip = 1, min(2, nexchpots(iorb))
     pfx = 'k1_'
     if (ip == 2) pfx = 'k2_'
     in1 = ins1(iorb,ip); in2 = ins2(iorb,ip); idel = deltam4pot(iorb,ip)
     isym = isyms(iorb,ip); ibexp = ibexcp(iorb,ip)
     do i = 1, mxsize
        rhs(i) = orbptr(i1b(in1)+i-1)*orbptr(i1b(in2)+i-1)*supplptr(i4b(10)+i-1)
        lhs(i) = supplptr(i4b(8)+i-1) + dble(idel*idel)*supplptr(i4b(4)+i-1)
     enddo
     wk = 0
     call putin(nni, mxnmu, isym, exchptr(ibexp:), wk)
     call dumpd(pfx//'win', mxsize8, wk)
     call dumpd(pfx//'lhs', mxsize, lhs)
     call dumpd(pfx//'rhs', mxsize, rhs)
     call dumpi(pfx//'isym', 1, (/ isym /))
     omega = ovfexch; omega1 = 1.0_PREC - omega
     maxsor2 = maxsorpot(2)
     wk0 = wk
     call mcsor(isym, wk0, lhs, rhs, supplptr(i4b(2):), supplptr(i4b(3):), &
          sorptr(iadext(1):), sorptr(iadnor(1):), sorptr(iadex1(1):), sorptr(iadex2(1):), sorptr(iadex3(1):))
     call dumpd(pfx//'wout_mcsor', mxsize8, wk0)
     nsw = maxsor2
     maxsor2 = 1
     wk0 = wk
     call mcsor(isym, wk0, lhs, rhs, supplptr(i4b(2):), supplptr(i4b(3):), &
          sorptr(iadext(1):), sorptr(iadnor(1):), sorptr(iadex1(1):), sorptr(iadex2(1):), sorptr(iadex3(1):))
     call dumpd(pfx//'wout_mcsor1', mxsize8, wk0)
     maxsor2 = nsw
     wk0 = wk
     call sor(isym, wk0, lhs, rhs, supplptr(i4b(2):), supplptr(i4b(3):), &
          sorptr(iadext(1):), sorptr(iadnor(1):), sorptr(iadex1(1):), sorptr(iadex2(1):), sorptr(iadex3(1):))
     call dumpd(pfx//'wout_sor', mxsize8, wk0)
     ! the backward sweep of the altsweeps option (doSCF.F90:403-411)
     isstart = ngrid1; isstop = 1; isstep = -1
     wk0 = wk
     call sor(isym, wk0, lhs, rhs, supplptr(i4b(2):), supplptr(i4b(3):), &
          sorptr(iadext(1):), sorptr(iadnor(1):), sorptr(iadex1(1):), sorptr(iadex2(1):), sorptr(iadex3(1):))
     call dumpd(pfx//'wout_sor_bwd', mxsize8, wk0)
     isstart = 1; isstop = ngrid1; isstep = 1
     xv = 0; xv(1) = omega; xv(2) = dble(maxsor2)
     call dumpd(pfx//'scal', 8, xv)
  enddo

  ! the relaxation driver of the potentials, then of the orbital
  if (nexchpots(iorb) /= 0) then
     if (lpotmcsor) then
        call coulExchMCSOR(iorb)
     else
        call coulExchSOR(iorb)
     endif
  endif
  call dumpd('exch_relaxed', int(length3), exchptr)

  ! Fock operator of iorb on the relaxed potentials (fock1 = scratch slot 1,
  ! fock2 = scratch slot 2)
  call fockHF(iorb)
  call dumpd('fock1', mxsize, scratchptr(1:))
  call dumpd('fock2', mxsize, scratchptr(mxsize8+1:))
  call dumpd('coulombptr', norb*mxsize, coulombptr)
  call dumpd('exchangeptr', norb*mxsize, exchangeptr)

  if (lorbmcsor) then
     call orbMCSOR(iorb)
  else
     call orbSOR(iorb)
  endif
  call dumpd('orb_relaxed', norb*mxsize, orbptr)

  call norm(iorb)
  call dumpd('orb_norm', norb*mxsize, orbptr)
  call dumpd('norm_1', maxorb, orbNorm)
  if (lhomonucl .and. .not. lbreakCi) call setci(iorb, orbptr)
  if (norb > 1) call ortho(iorb)
  call dumpd('orb_ortho', norb*mxsize, orbptr)

  ! one-electron integral <i|T+V|i> and the orbital energy
  allocate(wk1(mxsize8), wk2(mxsize8), wk3(mxsize8))
  wk0 = 0; wk1 = 0; wk2 = 0; wk3 = 0
  woneel = oneelij(iorb, iorb, orbptr, supplptr(i4b(4):), supplptr(i4b(5):), supplptr(i4b(13):), &
       wk0, wk1, wk2, wk3)
  call EaHF(iorb)
  call EabHF(iorb)
  call dumpd('ee_1', maxorb*maxorb, ee)
  xv = 0; xv(1) = woneel
  call dumpd('oneel_1', 8, xv)

  ! total energy and multipole moments on the state after this step
  itouch = 1
  call mpoleMoments
  call dumpd('cmulti_1', maxorb*maxmpole, cmulti)
  call dumpd('excdi_1', maxorb*(maxorb+1), excdi)
  call dumpd('excqu_1', maxorb*(maxorb+1), excqu)
  call dumpd('excoc_1', maxorb*(maxorb+1), excoc)
  call dumpd('exche_1', maxorb*(maxorb+1), exche)
  call etotalHF
  xv = 0; xv(1) = etot; xv(2) = vkt; xv(3) = vnt; xv(4) = evt; xv(5) = encoul; xv(6) = enexch
  call dumpd('etot_1', 8, xv)

contains

  subroutine dumpd(name, cnt, a)
    character(len=*), intent(in) :: name
    integer (KIND=IPREC), intent(in) :: cnt
    real (PREC), intent(in) :: a(*)
    open(newunit=iu, file=trim(ddir)//'/'//name//'.f64', access='stream', form='unformatted', &
         status='replace')
    write(iu) a(1:cnt)
    close(iu)
  end subroutine dumpd

  subroutine dumpi(name, cnt, a)
    character(len=*), intent(in) :: name
    integer (KIND=IPREC), intent(in) :: cnt
    integer (KIND=IPREC), intent(in) :: a(*)
    open(newunit=iu, file=trim(ddir)//'/'//name//'.i32', access='stream', form='unformatted', &
         status='replace')
    write(iu) a(1:cnt)
    close(iu)
  end subroutine dumpi

end program ref_harness
