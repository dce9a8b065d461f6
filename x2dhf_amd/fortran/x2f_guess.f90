! Starting orbitals and potentials on the host (set-up once per job):
!   hydrogenic LCAO orbitals     initHyd, hydrogenOrbA/B  initOrbitalsPotentials.F90:714-880
!   Thomas-Fermi Coulomb start   initCoulomb, pottf       coulombExchangePotentials.f90:14-71,222-239
!   1/r exchange start           initExchange             coulombExchangePotentials.f90:74-147
! Potentials are stored premultiplied by f4 = R xi/2, as the relaxation
! solves for them.
module x2f_guess
  use, intrinsic :: iso_c_binding
  use x2f_grid
  use x2f_case
  implicit none
  private
  public :: guess_orbitals, guess_coulomb, guess_exchange

  integer, parameter :: dp = c_double
  real(dp), parameter :: pii = 3.1415926535897932384626433832795_dp

contains

  real(dp) function factor(n)
    integer, intent(in) :: n
    integer :: i
    factor = 1.0_dp
    do i = 1, n
       factor = dble(i)*factor
    enddo
  end function factor

  ! confluent hypergeometric 1F1(np; lp; x) for np <= 0 (a polynomial)
  real(dp) function hypg1f1(np, lp, x)
    integer, intent(in) :: np, lp
    real(dp), intent(in) :: x
    real(dp) :: t
    integer :: i
    hypg1f1 = 1.0_dp
    if (np == 0) return
    t = dble(np)*x/dble(lp)
    hypg1f1 = hypg1f1 + t
    do i = 2, -np
       t = t*x*dble(np+i-1)/dble(lp+i-1)/dble(i)
       hypg1f1 = hypg1f1 + t
    enddo
  end function hypg1f1

  ! Gauss hypergeometric 2F1(na, nb; nc; x) for na <= 0 (a polynomial)
  real(dp) function hypg2f1(na, nb, nc, x)
    integer, intent(in) :: na, nb, nc
    real(dp), intent(in) :: x
    real(dp) :: a, b, c, t
    integer :: i
    hypg2f1 = 1.0_dp
    if (na == 0) return
    a = dble(na); b = dble(nb); c = dble(nc)
    t = a*b*x/c
    hypg2f1 = hypg2f1 + t
    do i = 2, -na
       t = t*x*(a+dble(i-1))*(b+dble(i-1))/(c+dble(i-1))/dble(i)
       hypg2f1 = hypg2f1 + t
    enddo
  end function hypg2f1

  ! associated Legendre function in the reference's normalisation and sign
  ! (plegendg, utils.f90:315-335; the prefactor evaluates to -1 for m /= 0)
  real(dp) function plegendg(l, m, c)
    integer, intent(in) :: l, m
    real(dp), intent(in) :: c
    real(dp) :: ct, fn0
    ct = abs(1.0_dp - c*c)
    if (m == 0) then
       plegendg = hypg2f1(-l, l+1, 1, (1.0_dp - c)/2.0_dp)
    elseif (ct < epsilon(1.0_dp)) then
       plegendg = 0.0_dp
    else
       fn0 = -1.0_dp*factor(l+m)/factor(l-m)/factor(m)/(2.0_dp**dble(m))
       plegendg = hypg2f1(m-l, m+l+1, m+1, (1.0_dp - c)/2.0_dp)*fn0*ct**(dble(m)/2.0_dp)
    endif
  end function plegendg

  real(dp) function plaguer(n, l, z, r)
    integer, intent(in) :: n, l
    real(dp), intent(in) :: z, r
    real(dp) :: x
    x = z*r/dble(n)
    plaguer = r**dble(l)*exp(-x)*hypg1f1(-(n-l-1), 2*l+2, x+x)
  end function plaguer

  ! hydrogenic function (n,l,m,z) centred on A (side = 1, nu = 0) or B (side = 2)
  subroutine hydrogenic(g, side, n, l, m, z, orb)
    type(grid_t), intent(in) :: g
    integer, intent(in) :: side, n, l, m
    real(dp), intent(in) :: z
    real(dp), intent(out) :: orb(g%nni*g%nmu)
    real(dp) :: fn, shn, zz, rt, ct, r
    integer :: in, imu
    r = g%r
    fn = (2.0_dp*z/dble(n))**(3.0_dp/2.0_dp + dble(l))*sqrt(factor(n+l)/(2.0_dp*dble(n)*factor(n-l-1))) &
         /factor(2*l+1)
    if (m == 0) then
       shn = 1.0_dp/sqrt(4.0_dp*pii)*sqrt((2.0_dp*l + 1.0_dp)*factor(l-m)/factor(l+m))
    else
       shn = (-1.0_dp)**dble(m)/sqrt(4.0_dp*pii)*sqrt((2.0_dp*l + 1.0_dp)*factor(l-m)/factor(l+m))
    endif
    do in = 1, g%nni
       do imu = 1, g%nmu
          zz = (r/2.0_dp)*g%vxi(imu)*g%veta(in)
          if (side == 1) then
             rt = (r/2.0_dp)*(g%vxi(imu) + g%veta(in))
             if (rt < epsilon(1.0_dp)) then
                ct = -1.0_dp
             else
                ct = max(-1.0_dp, min(1.0_dp, (zz + r/2.0_dp)/rt))
             endif
          else
             rt = (r/2.0_dp)*(g%vxi(imu) - g%veta(in))
             if (rt < epsilon(1.0_dp)) then
                ct = -1.0_dp
             else
                ct = max(-1.0_dp, min(1.0_dp, -(zz - r/2.0_dp)/rt))
             endif
          endif
          orb((imu-1)*g%nni + in) = fn*plaguer(n, l, z, rt)*shn*plegendg(l, m, ct)
       enddo
    enddo
  end subroutine hydrogenic

  ! LCAO orbitals (unnormalised, as initHyd leaves them)
  subroutine guess_orbitals(g, c, orbs)
    type(grid_t), intent(in) :: g
    type(case_t), intent(in) :: c
    real(dp), intent(out) :: orbs(g%nni*g%nmu, c%norb)
    real(dp), allocatable :: tmp(:)
    integer :: i
    allocate(tmp(g%nni*g%nmu))
    orbs = 0.0_dp
    do i = 1, c%norb
       if (abs(c%c1(i)) > epsilon(1.0_dp)) then
          call hydrogenic(g, 1, c%n1(i), c%l1(i), c%m(i), c%za(i), tmp)
          orbs(:,i) = orbs(:,i) + c%c1(i)*tmp
       endif
       if (abs(c%c2(i)) > epsilon(1.0_dp)) then
          call hydrogenic(g, 2, c%n2(i), c%l2(i), c%m(i), c%zb(i), tmp)
          orbs(:,i) = orbs(:,i) + c%c2(i)*tmp
       endif
    enddo
    deallocate(tmp)
  end subroutine guess_orbitals

  ! Thomas-Fermi screened potential of a unit charge at distance
  ! rab = R (xi + eta)/2 (pottf)
  real(dp) function pottf(r, eta, xi)
    real(dp), intent(in) :: r, eta, xi
    real(dp) :: rab, t, w
    pottf = 0.0_dp
    if (abs(xi + eta) < epsilon(1.0_dp)) return
    rab = r*(xi + eta)/2.0_dp
    w = sqrt(rab/0.885340_dp)
    t = w*(0.6011200_dp*w + 1.8106100_dp) + 1.0_dp
    w = w*(w*(w*(w*(0.0479300_dp*w + 0.2146500_dp) + 0.7711200_dp) + 1.3951500_dp) + 1.8106100_dp) + 1.0_dp
    pottf = (1.0_dp - (t/w)*(t/w))/rab
  end function pottf

  ! starting Coulomb potential of orbital i (times f4)
  subroutine guess_coulomb(g, c, i, pot)
    type(grid_t), intent(in) :: g
    type(case_t), intent(in) :: c
    integer, intent(in) :: i
    real(dp), intent(out) :: pot(g%nni*g%nmu)
    integer :: in, imu, k
    do in = 1, g%nni
       do imu = 1, g%nmu
          k = (imu-1)*g%nni + in
          if (imu == 1 .and. (in == 1 .or. in == g%nni)) then
             pot(k) = 0.0_dp
          elseif (nint(c%z1) /= 0 .and. nint(c%z2) /= 0) then
             pot(k) = -pottf(g%r, g%veta(in), g%vxi(imu))*abs(c%c1(i)) &
                  - pottf(g%r, -g%veta(in), g%vxi(imu))*abs(c%c2(i))
          else
             pot(k) = -pottf(g%r, g%veta(in), g%vxi(imu))*abs(c%c1(i))
          endif
       enddo
    enddo
    pot = g%suppl(:,9)*pot
  end subroutine guess_coulomb

  ! starting exchange potential of the pair (i, j), i <= j (times f4)
  subroutine guess_exchange(g, c, i, pot)
    type(grid_t), intent(in) :: g
    type(case_t), intent(in) :: c
    integer, intent(in) :: i
    real(dp), intent(out) :: pot(g%nni*g%nmu)
    integer :: in, imu, k
    real(dp) :: ra1, ra2
    pot = 0.0_dp
    do imu = 2, g%nmu
       do in = 2, g%nni - 1
          k = (imu-1)*g%nni + in
          ra1 = (g%r/2.0_dp)*(g%vxi(imu) + g%veta(in))
          ra2 = (g%r/2.0_dp)*(g%vxi(imu) - g%veta(in))
          if (ra1 > 1.0e-8_dp) then
             pot(k) = c%c1(i)/ra1
          else
             pot(k) = 0.0_dp
          endif
          if (ra2 > 1.0e-8_dp) pot(k) = pot(k) + c%c2(i)/ra2
       enddo
    enddo
    pot = g%suppl(:,9)*pot
  end subroutine guess_exchange
end module x2f_guess
