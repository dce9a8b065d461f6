"""CPU restatement of the starting functions of a job -- test infrastructure
only (tests/ may import it to check the device kernels of
x2dhf_amd/csrc/x2gpu_guess.hip; the product never does).

  hydrogenic LCAO orbitals  initHyd, hydrogenOrbA/B  initOrbitalsPotentials.F90:714-880
                            (plaguer / plegendg / hypg1f1 / hypg2f1, utils.f90:315-335)
  Thomas-Fermi Coulomb      initCoulomb, pottf       coulombExchangePotentials.f90:14-71, 222-239
  1/r exchange              initExchange             coulombExchangePotentials.f90:74-147

Functions take the grid vectors vxi (nmu) and veta (nni) and return nni*nmu
arrays, nu fastest.  Pinned in tests/test_guess.py by closed forms: the 1s, 2s
and 2p hydrogen functions, their unit norm under the grid quadrature, and the
Thomas-Fermi potential's 1/r tail.
"""
from math import factorial, pi, sqrt

import numpy as np

EPS = np.finfo(float).eps


def _hyp1f1(np_, lp, x):
    s = np.ones_like(x)
    if np_ == 0:
        return s
    t = np_ * x / lp
    s = s + t
    for i in range(2, -np_ + 1):
        t = t * x * (np_ + i - 1) / (lp + i - 1) / i
        s = s + t
    return s


def _hyp2f1(na, nb, nc, x):
    s = np.ones_like(x)
    if na == 0:
        return s
    t = na * nb * x / nc
    s = s + t
    for i in range(2, -na + 1):
        t = t * x * (na + i - 1) * (nb + i - 1) / (nc + i - 1) / i
        s = s + t
    return s


def legendre(l, m, c):
    """Associated Legendre function in the reference's normalisation and sign."""
    ct = np.abs(1.0 - c * c)
    if m == 0:
        return _hyp2f1(-l, l + 1, 1, (1.0 - c) / 2.0)
    fn0 = -1.0 * factorial(l + m) / factorial(l - m) / factorial(m) / 2.0 ** m
    out = _hyp2f1(m - l, m + l + 1, m + 1, (1.0 - c) / 2.0) * fn0 * ct ** (m / 2.0)
    return np.where(ct < EPS, 0.0, out)


def laguerre(n, l, z, r):
    x = z * r / n
    return r ** l * np.exp(-x) * _hyp1f1(-(n - l - 1), 2 * l + 2, x + x)


def hydrogenic(vxi, veta, r, side, n, l, m, z):
    """Hydrogenic function (n, l, m; z) on centre A (side 0) or B (side 1)."""
    fn = (2.0 * z / n) ** (1.5 + l) * sqrt(factorial(n + l) / (2.0 * n * factorial(n - l - 1))) / factorial(2 * l + 1)
    shn = (1.0 if m == 0 else (-1.0) ** m) / sqrt(4.0 * pi) * sqrt((2 * l + 1) * factorial(l - m) / factorial(l + m))
    xi, eta = np.meshgrid(vxi, veta, indexing="ij")      # (nmu, nni): nu fastest when raveled
    zz = (r / 2.0) * xi * eta
    if side == 0:
        rt = (r / 2.0) * (xi + eta)
        ct = np.where(rt < EPS, -1.0, np.clip((zz + r / 2.0) / np.where(rt < EPS, 1.0, rt), -1.0, 1.0))
    else:
        rt = (r / 2.0) * (xi - eta)
        ct = np.where(rt < EPS, -1.0, np.clip(-(zz - r / 2.0) / np.where(rt < EPS, 1.0, rt), -1.0, 1.0))
    return (fn * laguerre(n, l, z, rt) * shn * legendre(l, m, ct)).ravel()


def lcao(vxi, veta, r, m, c1, n1, l1, za, c2, n2, l2, zb):
    out = np.zeros(len(vxi) * len(veta))
    if abs(c1) > EPS:
        out += c1 * hydrogenic(vxi, veta, r, 0, n1, l1, m, za)
    if abs(c2) > EPS:
        out += c2 * hydrogenic(vxi, veta, r, 1, n2, l2, m, zb)
    return out


def pot_tf(r, eta, xi):
    """Thomas-Fermi screened potential of a unit charge at R (xi + eta) / 2."""
    rab = r * (xi + eta) / 2.0
    safe = np.where(np.abs(xi + eta) < EPS, 1.0, rab)
    w = np.sqrt(safe / 0.885340)
    t = w * (0.6011200 * w + 1.8106100) + 1.0
    w = w * (w * (w * (w * (0.0479300 * w + 0.2146500) + 0.7711200) + 1.3951500) + 1.8106100) + 1.0
    return np.where(np.abs(xi + eta) < EPS, 0.0, (1.0 - (t / w) * (t / w)) / safe)


def coulomb(vxi, veta, r, f4, c1, c2, two_centres):
    xi, eta = np.meshgrid(vxi, veta, indexing="ij")
    v = -pot_tf(r, eta, xi) * abs(c1)
    if two_centres:
        v = v - pot_tf(r, -eta, xi) * abs(c2)
    v[0, 0] = 0.0
    v[0, -1] = 0.0
    return f4 * v.ravel()


def exchange(vxi, veta, r, f4, c1, c2):
    xi, eta = np.meshgrid(vxi, veta, indexing="ij")
    ra1 = (r / 2.0) * (xi + eta)
    ra2 = (r / 2.0) * (xi - eta)
    v = np.where(ra1 > 1e-8, c1 / np.where(ra1 > 1e-8, ra1, 1.0), 0.0)
    v = v + np.where(ra2 > 1e-8, c2 / np.where(ra2 > 1e-8, ra2, 1.0), 0.0)
    v[0, :] = 0.0
    v[:, 0] = 0.0
    v[:, -1] = 0.0
    return f4 * v.ravel()
