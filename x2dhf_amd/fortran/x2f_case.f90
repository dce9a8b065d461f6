! The molecular case of one SCF job and its exchange bookkeeping, restated
! from the reference for closed-shell configurations:
!   orbital symmetries        initCBlocks  initCBAllocArrays.F90:896-987
!   exchange pairs / slots    initAddr     initCBAllocArrays.F90:1231-1273
!   potentials per orbital    initAddr     initCBAllocArrays.F90:1353-1444
!   exchange weights gec      initExWeights initCBAllocArrays.F90:3303-3432
!   exchange counts of E_HF   excont/exint exchContribs.f90:13-154
!   exchange moment list      initLP       initCBAllocArrays.F90:1596-1627
! Orbital i carries magnetic quantum number m(i) (0 sigma, 1 pi, 2 delta),
! g/u label gu(i) (+1 g, -1 u, 0 none), occupation occ(i) and the LCAO guess
! c1*phi_A(n1,l1,m,Z1) + c2*phi_B(n2,l2,m,Z2).  Orbitals are listed from the
! highest to the lowest, as in the reference's input; the SCF relaxes them in
! reverse order.
module x2f_case
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: case_t, MAXO, case_setup, pair_index

  integer, parameter :: dp = c_double
  integer, parameter :: MAXO = 16           ! orbitals per job
  integer, parameter :: MAXP = MAXO*(MAXO+1) ! exchange slots (two per pi-pi pair)

  type :: case_t
     integer :: norb = 0, nel = 0
     real(dp) :: z1 = 0, z2 = 0, r = 0
     logical :: homonuclear = .false.
     integer :: m(MAXO) = 0, gu(MAXO) = 0
     real(dp) :: occ(MAXO) = 0
     ! LCAO guess
     real(dp) :: c1(MAXO) = 0, c2(MAXO) = 0, za(MAXO) = 0, zb(MAXO) = 0
     integer :: n1(MAXO) = 0, l1(MAXO) = 0, n2(MAXO) = 0, l2(MAXO) = 0
     ! derived symmetries
     integer :: isym(MAXO) = 1        ! parity of the orbital across the axis (+1: m even)
     integer :: ihomo(MAXO) = 1       ! g/u sign used by setCi
     integer :: ige(MAXO) = 2         ! 1: u, 2: g or no label
     ! exchange pairs: k = i + j(j-1)/2 (i <= j); ilc(k) = 0, 1 or 2 slots;
     ! slot(k) and slot2(k) index the device exchange arrays
     integer :: k2(MAXO,MAXO) = 0
     integer :: ilc(MAXP) = 0, slot(MAXP) = 0, slot2(MAXP) = 0
     integer :: nslots = 0
     real(dp) :: gec(MAXO,MAXO) = 0, gec2(MAXO,MAXO) = 0
     ! potentials relaxed with orbital i: exchange with i..norb, then Coulomb of i
     integer :: npot(MAXO) = 0
     integer :: pin1(MAXO,2*MAXO+1) = 0, pin2(MAXO,2*MAXO+1) = 0, pdm(MAXO,2*MAXO+1) = 0
     integer :: psym(MAXO,2*MAXO+1) = 1, pslot(MAXO,2*MAXO+1) = 0   ! pslot = 0: Coulomb
     ! exchange moment list (pair, Delta m, slot)
     integer :: nmm = 0
     integer :: mm1(MAXP) = 0, mm2(MAXP) = 0, mmdel(MAXP) = 0, mmslot(MAXP) = 0
     ! exchange counts for the total energy
     real(dp) :: ox1(MAXO,MAXO) = 0, ox2(MAXO,MAXO) = 0, oxself(MAXO) = 0
  end type case_t

contains

  integer function pair_index(i, j)
    integer, intent(in) :: i, j
    integer :: a, b
    a = min(i, j); b = max(i, j)
    pair_index = a + b*(b-1)/2
  end function pair_index

  ! Fill the derived fields of c from norb, m, gu, occ, z1, z2, r.
  ! Returns 0, or a positive code for an unsupported case:
  ! 1 = too many orbitals, 2 = open shell, 3 = |m| > 2.
  integer function case_setup(c) result(ierr)
    type(case_t), intent(inout) :: c
    integer :: i, j, k, n, ix, jx, ii, jj, iloop(MAXO), iqm(4), spn(4)
    real(dp) :: full
    ierr = 0
    if (c%norb < 1 .or. c%norb > MAXO) then
       ierr = 1; return
    endif
    c%nel = 0
    do i = 1, c%norb
       if (c%m(i) < 0 .or. c%m(i) > 2) then
          ierr = 3; return
       endif
       full = 2.0_dp
       if (c%m(i) > 0) full = 4.0_dp
       if (abs(c%occ(i) - full) > 1.0e-12_dp) then
          ierr = 2; return
       endif
       c%nel = c%nel + nint(c%occ(i))
       if (mod(c%m(i), 2) == 0) then
          c%isym(i) = 1
       else
          c%isym(i) = -1
       endif
       ! -1 for u-sigma, u-delta, g-pi (initCBAllocArrays.F90:981-987)
       c%ihomo(i) = 1
       if (c%gu(i) == -1 .and. mod(c%m(i), 2) == 0) c%ihomo(i) = -1
       if (c%gu(i) == 1 .and. mod(c%m(i), 2) == 1) c%ihomo(i) = -1
       c%ige(i) = 2
       if (c%gu(i) == -1) c%ige(i) = 1
    enddo
    c%homonuclear = abs(c%z1 - c%z2) < 1.0e-6_dp

    ! exchange pairs and their device slots (initAddr)
    c%k2 = 0; c%ilc = 0; c%slot = 0; c%slot2 = 0
    n = 0
    do j = 1, c%norb
       do i = 1, j
          k = i + j*(j-1)/2
          if (i == j .and. c%m(i) == 0) cycle
          c%ilc(k) = 1
          c%k2(i,j) = k; c%k2(j,i) = k
          n = n + 1; c%slot(k) = n
          if (i /= j .and. c%m(i) /= 0 .and. c%m(j) /= 0) then
             c%ilc(k) = 2
             n = n + 1; c%slot2(k) = n
          endif
       enddo
    enddo
    c%nslots = n

    ! potentials relaxed with each orbital (initAddr, lfastexch and lcoulexch)
    do i = 1, c%norb
       n = 0
       do j = i, c%norb
          if (i == j .and. c%m(i) == 0) cycle
          k = pair_index(i, j)
          if (c%ilc(k) == 0) cycle
          n = n + 1
          c%pin1(i,n) = min(i,j); c%pin2(i,n) = max(i,j)
          c%pdm(i,n) = abs(c%m(i) - c%m(j))
          if (i == j) c%pdm(i,n) = 2*c%m(i)
          c%psym(i,n) = merge(1, -1, mod(c%pdm(i,n), 2) == 0)
          c%pslot(i,n) = c%slot(k)
          if (c%ilc(k) == 2) then
             n = n + 1
             c%pin1(i,n) = min(i,j); c%pin2(i,n) = max(i,j)
             c%pdm(i,n) = c%m(i) + c%m(j)
             c%psym(i,n) = merge(1, -1, mod(c%pdm(i,n), 2) == 0)
             c%pslot(i,n) = c%slot2(k)
          endif
       enddo
       n = n + 1
       c%pin1(i,n) = i; c%pin2(i,n) = i; c%pdm(i,n) = 0; c%psym(i,n) = 1; c%pslot(i,n) = 0
       c%npot(i) = n
    enddo

    ! exchange moments: every pair with a slot, Delta m as above (initLP)
    n = 0
    do i = 1, c%norb
       do j = i, c%norb
          if (i == j .and. c%m(i) == 0) cycle
          k = pair_index(i, j)
          n = n + 1
          c%mm1(n) = i; c%mm2(n) = j
          c%mmdel(n) = abs(c%m(j) - c%m(i))
          if (i == j) c%mmdel(n) = 2*c%m(i)
          c%mmslot(n) = c%slot(k)
          if (c%ilc(k) == 2) then
             n = n + 1
             c%mm1(n) = i; c%mm2(n) = j; c%mmdel(n) = c%m(i) + c%m(j); c%mmslot(n) = c%slot2(k)
          endif
       enddo
    enddo
    c%nmm = n

    ! exchange weights (initExWeights) for the closed-shell spin slots:
    ! slots 1,2 = alpha,beta with lambda +m; slots 3,4 = alpha,beta with -m
    spn = [1, 2, 1, 2]
    iqm = [1, 1, -1, -1]
    do i = 1, c%norb
       iloop(i) = 2
       if (c%m(i) > 0) iloop(i) = 4
    enddo
    c%gec = 0; c%gec2 = 0
    do i = 1, c%norb
       do ix = 1, iloop(i)
          do j = 1, c%norb
             do jx = 1, iloop(j)
                if (spn(ix) /= spn(jx)) cycle
                if (abs(c%m(i) - c%m(j)) == abs(c%m(i) + c%m(j))) then
                   if (i /= j) c%gec(i,j) = c%gec(i,j) + 1.0_dp
                else
                   if (i == j .and. iqm(ix) /= iqm(jx)) c%gec(i,j) = c%gec(i,j) + 1.0_dp
                   if (i /= j .and. iqm(ix) == iqm(jx)) c%gec(i,j) = c%gec(i,j) + 1.0_dp
                   if (i /= j .and. iqm(ix) /= iqm(jx)) c%gec2(i,j) = c%gec2(i,j) + 1.0_dp
                endif
             enddo
          enddo
       enddo
       c%gec(i,:) = c%gec(i,:)/dble(iloop(i))
       c%gec2(i,:) = c%gec2(i,:)/dble(iloop(i))
    enddo

    ! exchange counts of the total energy (excont / exint)
    c%ox1 = 0; c%ox2 = 0; c%oxself = 0
    do i = 1, c%norb
       if (c%m(i) > 0) then
          ! within a non-sigma shell: lambda+ with lambda- of equal spin
          do ii = 1, 2
             do jj = 3, 4
                if (spn(ii) == spn(jj)) c%oxself(i) = c%oxself(i) + 1.0_dp
             enddo
          enddo
       endif
       do j = i+1, c%norb
          if (c%m(i) == 0 .or. c%m(j) == 0) then
             do ii = 1, iloop(i)
                do jj = 1, iloop(j)
                   if (spn(ii) == spn(jj)) c%ox1(i,j) = c%ox1(i,j) + 1.0_dp
                enddo
             enddo
          else
             do ii = 1, 2
                do jj = 1, 2
                   if (spn(ii) == spn(jj)) c%ox1(i,j) = c%ox1(i,j) + 1.0_dp
                enddo
             enddo
             do ii = 3, 4
                do jj = 3, 4
                   if (spn(ii) == spn(jj)) c%ox1(i,j) = c%ox1(i,j) + 1.0_dp
                enddo
             enddo
             do ii = 1, 2
                do jj = 3, 4
                   if (spn(ii) == spn(jj)) c%ox2(i,j) = c%ox2(i,j) + 1.0_dp
                enddo
             enddo
             do ii = 3, 4
                do jj = 1, 2
                   if (spn(ii) == spn(jj)) c%ox2(i,j) = c%ox2(i,j) + 1.0_dp
                enddo
             enddo
          endif
       enddo
    enddo
  end function case_setup
end module x2f_case
