! Prolate-spheroidal (nu,mu) grid and the coefficient arrays of the
! relaxation path for point nuclei: the host-side counterpart of the
! reference's initCBlocks (initCBAllocArrays.F90:726-1070) and initSuppl
! (initCBAllocArrays.F90:1636-2051), plus the automatic overrelaxation
! parameters (setOmega, initCBAllocArrays.F90:636-650, 1010-1033).
!
! Conventions: nu index i = 1..nni, mu index j = 1..nmu, arrays column-major
! with nu fastest; xi = cosh(mu), eta = cos(nu).  The 14 coefficient arrays
! are kept in the reference's order (borb, bpot, d, e, f0, f1, f2, f3, f4, g,
! wjac1, wjac2, wgt1, wgt2) so the device library indexes them the same way.
module x2f_grid
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: grid_t, grid_build, grid_coef, grid_vec, omega_defaults

  integer, parameter :: dp = c_double
  real(dp), parameter :: pii = 3.1415926535897932384626433832795_dp

  type :: grid_t
     integer :: nni = 0, nmu = 0
     real(dp) :: z1 = 0, z2 = 0, r = 0, rinf = 0
     real(dp) :: hni = 0, hmu = 0, diag = 0
     real(dp) :: dmu1(4), dmu2(4), dni1(4), dni2(4), exeven(5)
     real(dp), allocatable :: vmu(:), vxi(:), vxisq(:), vxi1(:), vxi2(:)
     real(dp), allocatable :: vni(:), veta(:), vetasq(:), veta1(:), veta2(:)
     real(dp), allocatable :: suppl(:,:)      ! (nni*nmu, 14)
  end type grid_t

contains

  ! Build the grid of an (nni x nmu) mesh with practical infinity rinf for
  ! nuclei z1 (at nu = 0) and z2 (at nu = pi) a distance r apart.
  ! Returns 0, or 1 when the sizes do not fit the 7-point quadrature.
  integer function grid_build(g, nni, nmu, z1, z2, r, rinf) result(ierr)
    type(grid_t), intent(inout) :: g
    integer, intent(in) :: nni, nmu
    real(dp), intent(in) :: z1, z2, r, rinf
    real(dp), parameter :: aa1(9) = [3.0_dp, -32.0_dp, 168.0_dp, -672.0_dp, 0.0_dp, &
         672.0_dp, -168.0_dp, 32.0_dp, -3.0_dp]
    real(dp), parameter :: aa2(9) = [-9.0_dp, 128.0_dp, -1008.0_dp, 8064.0_dp, -14350.0_dp, &
         8064.0_dp, -1008.0_dp, 128.0_dp, -9.0_dp]
    real(dp) :: xmi0, w1, w2, w4, w5, w6, w7, w8, w9, w10, wj1, wj2, xrr, wmup
    real(dp), allocatable :: wmu(:), wni(:)
    integer :: i, j, k, n

    ierr = 0
    if (nni < 13 .or. nmu < 13 .or. mod(nni-1, 6) /= 0 .or. mod(nmu-1, 6) /= 0) then
       ierr = 1
       return
    endif
    g%nni = nni; g%nmu = nmu; g%z1 = z1; g%z2 = z2; g%r = r; g%rinf = rinf
    n = nni*nmu
    if (allocated(g%vmu)) deallocate(g%vmu, g%vxi, g%vxisq, g%vxi1, g%vxi2, g%vni, g%veta, &
         g%vetasq, g%veta1, g%veta2, g%suppl)
    allocate(g%vmu(nmu), g%vxi(nmu), g%vxisq(nmu), g%vxi1(nmu), g%vxi2(nmu))
    allocate(g%vni(nni), g%veta(nni), g%vetasq(nni), g%veta1(nni), g%veta2(nni))
    allocate(g%suppl(n, 14), wmu(nmu), wni(nni))

    ! steps: h_nu = pi/(nni-1); h_mu from the practical infinity
    ! (initCBAllocArrays.F90:813-822)
    g%hni = pii/dble(nni-1)
    xmi0 = 2.0_dp*rinf/r
    xmi0 = log(xmi0 + sqrt(xmi0*xmi0 - 1.0_dp))
    g%hmu = xmi0/dble(nmu-1)

    ! mu, xi tables: mu accumulated step by step (initCBAllocArrays.F90:826-844)
    g%vmu(1) = 0.0_dp
    g%vxi(1) = cosh(g%vmu(1))
    g%vxisq(1) = g%vxi(1)*g%vxi(1)
    g%vxi1(1) = sqrt(g%vxi(1)*g%vxi(1) - 1.0_dp)
    g%vxi2(1) = 0.0_dp
    do i = 2, nmu
       g%vmu(i) = g%vmu(i-1) + g%hmu
       g%vxi(i) = cosh(g%vmu(i))
       g%vxisq(i) = g%vxi(i)*g%vxi(i)
       g%vxi1(i) = sqrt(g%vxi(i)*g%vxi(i) - 1.0_dp)
       g%vxi2(i) = g%vxi(i)/g%vxi1(i)
    enddo
    ! nu, eta tables (initCBAllocArrays.F90:848-860)
    do i = 1, nni
       g%vni(i) = dble(i-1)*g%hni
       g%veta(i) = cos(g%vni(i))
       g%vetasq(i) = g%veta(i)*g%veta(i)
       g%veta1(i) = sqrt(1.0_dp - g%veta(i)*g%veta(i))
       if (g%veta1(i) < epsilon(1.0_dp)) then
          g%veta2(i) = 0.0_dp
       else
          g%veta2(i) = g%veta(i)/g%veta1(i)
       endif
    enddo

    ! coefficient arrays (initSuppl, initCBAllocArrays.F90:1755-1903)
    wj1 = -r*pii/2.0_dp
    wj2 = r*r*pii/2.0_dp
    xrr = 0.5_dp*r
    g%diag = -14350.0_dp/(5040.0_dp)*(1.0_dp/(g%hmu*g%hmu) + 1.0_dp/(g%hni*g%hni))
    do i = 1, nmu
       w1 = g%vxi1(i)
       w2 = g%vxi(i)*g%vxi(i)
       w9 = xrr*g%vxi(i)
       w4 = g%vxi2(i)
       if (i == 1) then
          w5 = 0.0_dp
          w6 = 0.0_dp
       else
          w5 = g%vxi(i)/g%vxi1(i) - 2.0_dp*g%vxi1(i)/g%vxi(i)
          w6 = 1.0_dp/(g%vxi1(i)*g%vxi1(i))
       endif
       w8 = -r*r*r*pii*g%vxi(i)/2.0_dp
       do j = 1, nni
          k = (i-1)*nni + j
          w10 = g%veta1(j)
          g%suppl(k, 1) = w4                                    ! borb
          g%suppl(k, 2) = w5                                    ! bpot
          g%suppl(k, 3) = g%veta2(j)                            ! d
          if (w10 < epsilon(1.0_dp)) then
             g%suppl(k, 4) = 0.0_dp                             ! e
          else
             g%suppl(k, 4) = -w6 - 1.0_dp/(w10*w10)
          endif
          w7 = w2 - g%veta(j)*g%veta(j)
          g%suppl(k, 6) = r*r*w7/2.0_dp                         ! f1
          g%suppl(k, 5) = r*(g%vxi(i)*(z1+z2) + g%veta(j)*(z2-z1))  ! f0, point nuclei
          if (i == 1) then
             g%suppl(k, 7) = 0.0_dp                             ! f2
          else
             g%suppl(k, 7) = -r*w7/g%vxi(i)
          endif
          g%suppl(k, 8) = -2.0_dp/w2 + g%diag                   ! f3 (+diag)
          g%suppl(k, 10) = w8*w7                                ! g
          g%suppl(k, 9) = w9                                    ! f4
          g%suppl(k, 11) = wj1*w1*w10                           ! wjac1
          g%suppl(k, 12) = wj2*w1*w10*(w1*w1 + w10*w10)/g%vxi(i) ! wjac2
       enddo
    enddo

    ! composite 7-point Newton-Cotes weights (initCBAllocArrays.F90:1917-1952)
    wmup = 0.0_dp
    w1 = g%hmu/140.0_dp
    wmu(1) = wmup + w1*41.0_dp
    do i = 2, nmu, 6
       wmu(i) = w1*216.0_dp
       wmu(i+1) = w1*27.0_dp
       wmu(i+2) = w1*272.0_dp
       wmu(i+3) = w1*27.0_dp
       wmu(i+4) = w1*216.0_dp
       if (i+5 == nmu) then
          wmu(i+5) = w1*41.0_dp
       else
          wmu(i+5) = 2.0_dp*w1*41.0_dp
       endif
    enddo
    w1 = g%hni/140.0_dp
    wni(1) = w1*41.0_dp
    do i = 2, nni, 6
       wni(i) = w1*216.0_dp
       wni(i+1) = w1*27.0_dp
       wni(i+2) = w1*272.0_dp
       wni(i+3) = w1*27.0_dp
       wni(i+4) = w1*216.0_dp
       if (i+5 == nni) then
          wni(i+5) = w1*41.0_dp
       else
          wni(i+5) = 2.0_dp*w1*41.0_dp
       endif
    enddo
    ! Jacobians times the mu weights, then the nu weights (initCBAllocArrays.F90:1955-1972)
    do i = 1, nmu
       w1 = wmu(i)
       do j = 1, nni
          k = (i-1)*nni + j
          g%suppl(k, 11) = w1*g%suppl(k, 11)
          g%suppl(k, 12) = w1*g%suppl(k, 12)
       enddo
    enddo
    do i = 1, nmu
       do j = 1, nni
          k = (i-1)*nni + j
          g%suppl(k, 13) = wni(j)*g%suppl(k, 11)
          g%suppl(k, 14) = wni(j)*g%suppl(k, 12)
       enddo
    enddo

    ! 9-point first/second derivative weights (initCBAllocArrays.F90:1979-1990)
    w1 = 1.0_dp/(840.0_dp)
    w2 = 1.0_dp/(5040.0_dp)
    do i = 1, 4
       g%dmu2(i) = w2*aa2(i)/(g%hmu*g%hmu)
       g%dmu1(i) = w1*aa1(i)/g%hmu
       g%dni2(i) = w2*aa2(i)/(g%hni*g%hni)
       g%dni1(i) = w1*aa1(i)/g%hni
    enddo
    ! boundary extrapolation weights of even functions (initCBAllocArrays.F90:2037-2041)
    g%exeven = [210.0_dp, -120.0_dp, 45.0_dp, -10.0_dp, 1.0_dp]/126.0_dp
    deallocate(wmu, wni)
  end function grid_build

  ! The 25-double coefficient block of x2g_grid_create.
  subroutine grid_coef(g, coef)
    type(grid_t), intent(in) :: g
    real(dp), intent(out) :: coef(25)
    coef(1:4) = g%dmu1; coef(5:8) = g%dmu2; coef(9:12) = g%dni1; coef(13:16) = g%dni2
    coef(17:21) = g%exeven
    coef(22) = g%diag; coef(23) = g%hni; coef(24) = g%hmu; coef(25) = g%r
  end subroutine grid_coef

  ! {vxi, vxisq, veta, vetasq} of x2g_grid_create.
  subroutine grid_vec(g, vec)
    type(grid_t), intent(in) :: g
    real(dp), intent(out) :: vec(2*g%nmu + 2*g%nni)
    vec(1:g%nmu) = g%vxi
    vec(g%nmu+1:2*g%nmu) = g%vxisq
    vec(2*g%nmu+1:2*g%nmu+g%nni) = g%veta
    vec(2*g%nmu+g%nni+1:) = g%vetasq
  end subroutine grid_vec

  ! Automatic overrelaxation parameters (setOmega, initCBAllocArrays.F90:636-650;
  ! orbital value initCBAllocArrays.F90:1018-1033 with the default Z correction).
  subroutine omega_defaults(nni, nmu, z1, z2, omega_orb, omega_pot)
    integer, intent(in) :: nni, nmu
    real(dp), intent(in) :: z1, z2
    real(dp), intent(out) :: omega_orb, omega_pot
    real(dp), parameter :: a = 0.603_dp, b = 0.79_dp, zfactor = 0.001_dp
    real(dp) :: rho
    rho = 0.5_dp*(cos(pii/dble(nmu)) + cos(pii/dble(nni)))
    omega_pot = 2.0_dp*a/(1.0_dp + sqrt(1.0_dp - rho*rho)) + b
    omega_orb = (omega_pot - 1.0_dp)*omega_pot - zfactor*max(z1, z2)
  end subroutine omega_defaults
end module x2f_grid
