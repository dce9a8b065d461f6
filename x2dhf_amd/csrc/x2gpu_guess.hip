// Starting orbitals and potentials on the device (set-up of a job):
//   hydrogenic LCAO orbitals     initHyd, hydrogenOrbA/B   initOrbitalsPotentials.F90:714-880
//   Thomas-Fermi Coulomb start   initCoulomb, pottf        coulombExchangePotentials.f90:14-71, 222-239
//   1/r exchange start           initExchange              coulombExchangePotentials.f90:74-147
// One thread per grid point; the orbital's normalisation constants (factorials)
// are formed on the host.  Potentials are stored premultiplied by f4 = R xi / 2,
// the form the relaxation solves for.
#include "x2gpu_internal.h"

namespace x2g {

struct HydDesc {
    double c[2], z[2];   // mixing coefficient and exponent on centres A, B
    double fn[2];        // radial normalisation (2z/n)^(3/2+l) sqrt((n+l)! / (2n (n-l-1)!)) / (2l+1)!
    double shn[2];       // angular normalisation, (-1)^m sqrt((2l+1)(l-m)! / (4 pi (l+m)!))
    double fn0[2];       // prefactor of the associated Legendre function for m > 0
    int n[2], l[2];
    int m;
};

// 1F1(np; lp; x) for np <= 0 (a polynomial of degree -np)
__device__ double hyp1f1(int np, int lp, double x)
{
    double s = 1.0;
    if (np == 0) return s;
    double t = (double)np * x / (double)lp;
    s += t;
    for (int i = 2; i <= -np; ++i) {
        t = t * x * (double)(np + i - 1) / (double)(lp + i - 1) / (double)i;
        s += t;
    }
    return s;
}

// 2F1(na, nb; nc; x) for na <= 0 (a polynomial of degree -na)
__device__ double hyp2f1(int na, int nb, int nc, double x)
{
    double s = 1.0;
    if (na == 0) return s;
    const double a = na, b = nb, c = nc;
    double t = a * b * x / c;
    s += t;
    for (int i = 2; i <= -na; ++i) {
        t = t * x * (a + (double)(i - 1)) * (b + (double)(i - 1)) / (c + (double)(i - 1)) / (double)i;
        s += t;
    }
    return s;
}

// associated Legendre function in the reference's normalisation and sign
// (plegendg, utils.f90:315-335)
__device__ double legendre(int l, int m, double c, double fn0)
{
    const double ct = fabs(1.0 - c * c);
    if (m == 0) return hyp2f1(-l, l + 1, 1, (1.0 - c) / 2.0);
    if (ct < 2.220446049250313e-16) return 0.0;
    return hyp2f1(m - l, m + l + 1, m + 1, (1.0 - c) / 2.0) * fn0 * pow(ct, (double)m / 2.0);
}

// r^l e^{-z r / n} 1F1(-(n-l-1); 2l+2; 2 z r / n)  (plaguer)
__device__ double laguerre(int n, int l, double z, double r)
{
    const double x = z * r / (double)n;
    double rl = 1.0;
    for (int k = 0; k < l; ++k) rl *= r;
    return rl * exp(-x) * hyp1f1(-(n - l - 1), 2 * l + 2, x + x);
}

__global__ void k_guess_orbital(double *orb, const double *vxi, const double *veta, int nni, int nmu, double r,
                                HydDesc h)
{
    const long tot = (long)nni * nmu;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int imu = (int)(e / nni), in = (int)(e - (long)imu * nni);
        const double xi = vxi[imu], eta = veta[in];
        const double zz = (r / 2.0) * xi * eta;
        double v = 0.0;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            if (!(fabs(h.c[side]) > 2.220446049250313e-16)) continue;
            double rt, ct;
            if (side == 0) {
                rt = (r / 2.0) * (xi + eta);
                ct = rt < 2.220446049250313e-16 ? -1.0 : fmax(-1.0, fmin(1.0, (zz + r / 2.0) / rt));
            } else {
                rt = (r / 2.0) * (xi - eta);
                ct = rt < 2.220446049250313e-16 ? -1.0 : fmax(-1.0, fmin(1.0, -(zz - r / 2.0) / rt));
            }
            v += h.c[side] * (h.fn[side] * laguerre(h.n[side], h.l[side], h.z[side], rt) * h.shn[side] *
                              legendre(h.l[side], h.m, ct, h.fn0[side]));
        }
        orb[e] = v;
    }
}

// Thomas-Fermi screened potential of a unit charge at distance R (xi + eta) / 2 (pottf)
__device__ double pot_tf(double r, double eta, double xi)
{
    if (fabs(xi + eta) < 2.220446049250313e-16) return 0.0;
    const double rab = r * (xi + eta) / 2.0;
    double w = sqrt(rab / 0.885340);
    const double t = w * (0.6011200 * w + 1.8106100) + 1.0;
    w = w * (w * (w * (w * (0.0479300 * w + 0.2146500) + 0.7711200) + 1.3951500) + 1.8106100) + 1.0;
    return (1.0 - (t / w) * (t / w)) / rab;
}

__global__ void k_guess_coulomb(double *pot, const double *vxi, const double *veta, const double *f4, int nni,
                                int nmu, double r, double c1, double c2, int two_centres)
{
    const long tot = (long)nni * nmu;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int imu = (int)(e / nni), in = (int)(e - (long)imu * nni);
        double v;
        if (imu == 0 && (in == 0 || in == nni - 1)) {
            v = 0.0;
        } else if (two_centres) {
            v = -pot_tf(r, veta[in], vxi[imu]) * fabs(c1) - pot_tf(r, -veta[in], vxi[imu]) * fabs(c2);
        } else {
            v = -pot_tf(r, veta[in], vxi[imu]) * fabs(c1);
        }
        pot[e] = f4[e] * v;
    }
}

__global__ void k_guess_exchange(double *pot, const double *vxi, const double *veta, const double *f4, int nni,
                                 int nmu, double r, double c1, double c2)
{
    const long tot = (long)nni * nmu;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int imu = (int)(e / nni), in = (int)(e - (long)imu * nni);
        double v = 0.0;
        if (imu >= 1 && in >= 1 && in <= nni - 2) {
            const double ra1 = (r / 2.0) * (vxi[imu] + veta[in]);
            const double ra2 = (r / 2.0) * (vxi[imu] - veta[in]);
            if (ra1 > 1.0e-8) v = c1 / ra1;
            if (ra2 > 1.0e-8) v += c2 / ra2;
        }
        pot[e] = f4[e] * v;
    }
}

static double factorial(int n)
{
    double f = 1.0;
    for (int i = 1; i <= n; ++i) f *= (double)i;
    return f;
}

static int guess_grid_x(const x2g_grid *g) { return (int)std::min<size_t>((g->n() + 255) / 256, 4096); }

}  // namespace x2g

using namespace x2g;

extern "C" int x2g_guess_orbital(x2g_grid *g, double *orb_d, const x2g_lcao *lc)
{
    X2G_CHECK(g && orb_d && lc, "x2g_guess_orbital: null argument");
    HydDesc h{};
    h.m = lc->m;
    const double cs[2] = {lc->c1, lc->c2}, zs[2] = {lc->za, lc->zb};
    const int ns[2] = {lc->n1, lc->n2}, ls[2] = {lc->l1, lc->l2};
    for (int s = 0; s < 2; ++s) {
        h.c[s] = cs[s];
        h.z[s] = zs[s];
        h.n[s] = ns[s];
        h.l[s] = ls[s];
        if (!(std::fabs(cs[s]) > 2.220446049250313e-16)) continue;
        const int n = ns[s], l = ls[s], m = lc->m;
        X2G_CHECK(n >= 1 && l >= 0 && l < n && m >= 0 && m <= l,
                  "x2g_guess_orbital: centre %c needs 0 <= m <= l < n (n=%d l=%d m=%d)", 'A' + s, n, l, m);
        h.fn[s] = std::pow(2.0 * zs[s] / (double)n, 1.5 + (double)l) *
                  std::sqrt(factorial(n + l) / (2.0 * (double)n * factorial(n - l - 1))) / factorial(2 * l + 1);
        const double pi = 3.1415926535897932384626433832795;
        h.shn[s] = (m == 0 ? 1.0 : std::pow(-1.0, (double)m)) / std::sqrt(4.0 * pi) *
                   std::sqrt((2.0 * l + 1.0) * factorial(l - m) / factorial(l + m));
        h.fn0[s] = -1.0 * factorial(l + m) / factorial(l - m) / factorial(m) / std::pow(2.0, (double)m);
    }
    x2g_ctx *ctx = g->ctx;
    X2G_LAUNCH(ctx, k_guess_orbital, dim3(guess_grid_x(g)), dim3(256), 0, orb_d, g->vxi, g->veta, g->nni, g->nmu,
               g->coef[24], h);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_guess_coulomb(x2g_grid *g, double *pot_d, double c1, double c2, int two_centres)
{
    X2G_CHECK(g && pot_d, "x2g_guess_coulomb: null argument");
    x2g_ctx *ctx = g->ctx;
    X2G_LAUNCH(ctx, k_guess_coulomb, dim3(guess_grid_x(g)), dim3(256), 0, pot_d, g->vxi, g->veta,
               g->suppl + 8 * g->n(), g->nni, g->nmu, g->coef[24], c1, c2, two_centres);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_guess_exchange(x2g_grid *g, double *pot_d, double c1, double c2)
{
    X2G_CHECK(g && pot_d, "x2g_guess_exchange: null argument");
    x2g_ctx *ctx = g->ctx;
    X2G_LAUNCH(ctx, k_guess_exchange, dim3(guess_grid_x(g)), dim3(256), 0, pot_d, g->vxi, g->veta,
               g->suppl + 8 * g->n(), g->nni, g->nmu, g->coef[24], c1, c2);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}
