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
  !
  ! The relaxation solves, in xi = cosh(mu), eta = cos(nu), the equations
  ! multiplied through by the metric factor r^2 (xi^2 - eta^2)/4, which
  ! leaves d2/dmu2 + d2/dnu2 with unit weight.  With q = xi^2 - eta^2 the
  ! 14 coefficient arrays are, point by point:
  !   borb   coth(mu)                         d/dmu of an orbital
  !   bpot   coth(mu) - 2 tanh(mu)            d/dmu of a potential
  !   d      cot(nu)                          d/dnu
  !   e      -1/sinh^2(mu) - 1/sin^2(nu)      the m^2 term
  !   f0     r (xi (z1+z2) + eta (z2-z1))     (z1/r1 + z2/r2) r^2 q/2
  !   f1     r^2 q/2                          the orbital energy's factor
  !   f2     -r q/xi
  !   f3     -2/xi^2 + the centre weight of both 9-point second differences
  !   f4     r xi/2
  !   g      -pi r^3 xi q/2                   the Poisson source's factor
  !   wjac1  -(pi r/2) sinh(mu) sin(nu)                      volume element
  !   wjac2  (pi r^2/2) sinh(mu) sin(nu) (sinh^2 mu + sin^2 nu)/xi
  ! wjac1, wjac2 carry the mu quadrature weights; wgt1, wgt2 are those times
  ! the nu weights.  On the axis row mu = 0, bpot and f2 are zero and e keeps
  ! only its nu term; where sin(nu) = 0, d and e are zero.  Every product is
  ! evaluated in the same order as the reference's initSuppl
  ! (initCBAllocArrays.F90:1755-1990), so the arrays agree with its to the
  ! last bit and the relaxation follows the same trajectory.
  integer function grid_build(g, nni, nmu, z1, z2, r, rinf) result(ierr)
    type(grid_t), intent(inout) :: g
    integer, intent(in) :: nni, nmu
    real(dp), intent(in) :: z1, z2, r, rinf
    ! 9-point central differences: points k-4..k-1 of the first (/840 h) and
    ! second (/5040 h^2) derivative; the centre of the second is -14350
    real(dp), parameter :: c1(4) = [3.0_dp, -32.0_dp, 168.0_dp, -672.0_dp]
    real(dp), parameter :: c2(4) = [-9.0_dp, 128.0_dp, -1008.0_dp, 8064.0_dp]
    real(dp) :: x
    real(dp), allocatable :: xi(:,:), eta(:,:), sh(:,:), sn(:,:), q(:,:), c(:,:), wmu(:), wni(:)
    integer :: i

    ierr = 0
    if (nni < 13 .or. nmu < 13 .or. mod(nni-1, 6) /= 0 .or. mod(nmu-1, 6) /= 0) then
       ierr = 1
       return
    endif
    g%nni = nni; g%nmu = nmu; g%z1 = z1; g%z2 = z2; g%r = r; g%rinf = rinf
    if (allocated(g%vmu)) deallocate(g%vmu, g%vxi, g%vxisq, g%vxi1, g%vxi2, g%vni, g%veta, &
         g%vetasq, g%veta1, g%veta2, g%suppl)
    allocate(g%vmu(nmu), g%vxi(nmu), g%vxisq(nmu), g%vxi1(nmu), g%vxi2(nmu))
    allocate(g%vni(nni), g%veta(nni), g%vetasq(nni), g%veta1(nni), g%veta2(nni))
    allocate(g%suppl(nni*nmu, 14))

    ! steps: h_nu = pi/(nni-1); the last mu node sits at the practical
    ! infinity, xi_max = 2 rinf/r (initCBAllocArrays.F90:813-822)
    g%hni = pii/dble(nni-1)
    x = 2.0_dp*rinf/r
    g%hmu = log(x + sqrt(x*x - 1.0_dp))/dble(nmu-1)

    ! axis tables (initCBAllocArrays.F90:826-860).  mu is accumulated node by
    ! node, as the reference's mesh is, not formed as (j-1)*h_mu.
    g%vmu(1) = 0.0_dp
    do i = 2, nmu
       g%vmu(i) = g%vmu(i-1) + g%hmu
    enddo
    g%vxi = cosh(g%vmu)
    g%vxisq = g%vxi*g%vxi
    g%vxi1 = sqrt(g%vxi*g%vxi - 1.0_dp)              ! sinh(mu)
    g%vxi2(1) = 0.0_dp
    g%vxi2(2:) = g%vxi(2:)/g%vxi1(2:)                 ! coth(mu)
    g%vni = [(dble(i-1)*g%hni, i = 1, nni)]
    g%veta = cos(g%vni)
    g%vetasq = g%veta*g%veta
    g%veta1 = sqrt(1.0_dp - g%veta*g%veta)           ! sin(nu)
    where (g%veta1 < epsilon(1.0_dp))
       g%veta2 = 0.0_dp
    elsewhere
       g%veta2 = g%veta/g%veta1                       ! cot(nu)
    end where

    ! the coefficient arrays as (nu, mu) fields
    xi = spread(g%vxi, 1, nni);   eta = spread(g%veta, 2, nmu)
    sh = spread(g%vxi1, 1, nni);  sn = spread(g%veta1, 2, nmu)
    q = xi*xi - eta*eta
    g%diag = -14350.0_dp/5040.0_dp*(1.0_dp/(g%hmu*g%hmu) + 1.0_dp/(g%hni*g%hni))
    allocate(c(nni, nmu))

    call put(1, spread(g%vxi2, 1, nni))
    c(:, 1) = 0.0_dp
    c(:, 2:) = xi(:, 2:)/sh(:, 2:) - 2.0_dp*sh(:, 2:)/xi(:, 2:)
    call put(2, c)
    call put(3, spread(g%veta2, 2, nmu))
    ! e: the 1/sinh^2 term is left out on the axis row
    c(:, 1) = 0.0_dp
    c(:, 2:) = 1.0_dp/(sh(:, 2:)*sh(:, 2:))
    where (sn < epsilon(1.0_dp))
       c = 0.0_dp
    elsewhere
       c = -c - 1.0_dp/(sn*sn)
    end where
    call put(4, c)
    call put(5, r*(xi*(z1+z2) + eta*(z2-z1)))
    call put(6, r*r*q/2.0_dp)
    c(:, 1) = 0.0_dp
    c(:, 2:) = -r*q(:, 2:)/xi(:, 2:)
    call put(7, c)
    call put(8, -2.0_dp/(xi*xi) + g%diag)
    call put(9, (0.5_dp*r)*xi)
    call put(10, (-r*r*r*pii*xi/2.0_dp)*q)

    ! volume elements with the composite 7-point Newton-Cotes weights
    ! (initCBAllocArrays.F90:1917-1972)
    wmu = newton_cotes7(nmu, g%hmu)
    wni = newton_cotes7(nni, g%hni)
    c = spread(wmu, 1, nni)*((-r*pii/2.0_dp)*sh*sn)
    call put(11, c)
    call put(13, spread(wni, 2, nmu)*c)
    c = spread(wmu, 1, nni)*((r*r*pii/2.0_dp)*sh*sn*(sh*sh + sn*sn)/xi)
    call put(12, c)
    call put(14, spread(wni, 2, nmu)*c)

    ! difference weights (initCBAllocArrays.F90:1979-1990)
    g%dmu1 = (1.0_dp/840.0_dp)*c1/g%hmu
    g%dmu2 = (1.0_dp/5040.0_dp)*c2/(g%hmu*g%hmu)
    g%dni1 = (1.0_dp/840.0_dp)*c1/g%hni
    g%dni2 = (1.0_dp/5040.0_dp)*c2/(g%hni*g%hni)
    ! boundary extrapolation weights of even functions (initCBAllocArrays.F90:2037-2041)
    g%exeven = [210.0_dp, -120.0_dp, 45.0_dp, -10.0_dp, 1.0_dp]/126.0_dp
  contains
    subroutine put(k, f)
      integer, intent(in) :: k
      real(dp), intent(in) :: f(:,:)
      g%suppl(:, k) = reshape(f, [nni*nmu])
    end subroutine put
  end function grid_build

  ! Weights of the composite 7-point closed Newton-Cotes rule, (41, 216, 27,
  ! 272, 27, 216, 41) h/140 per panel, over n = 6k+1 nodes spaced h: nodes
  ! shared by two panels take both panels' end weight.
  pure function newton_cotes7(n, h) result(w)
    integer, intent(in) :: n
    real(dp), intent(in) :: h
    real(dp) :: w(n)
    real(dp), parameter :: cycle6(0:5) = [82.0_dp, 216.0_dp, 27.0_dp, 272.0_dp, 27.0_dp, 216.0_dp]
    integer :: k
    do k = 1, n
       w(k) = (h/140.0_dp)*cycle6(mod(k-1, 6))
    enddo
    w(1) = (h/140.0_dp)*41.0_dp
    w(n) = w(1)
  end function newton_cotes7

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
