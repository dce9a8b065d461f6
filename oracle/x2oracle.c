/*
 * CPU restatement of the x2dhf relaxation kernels -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product path (x2dhf_amd) never calls it.
 *
 * Layout conventions (the reference's, inout.F90:23-82 and
 * initCBAllocArrays.F90:2446-2450):
 *   - a function f(nu,mu) on the nni x nmu grid is stored column-major with nu
 *     fastest: f[(mu-1)*nni + (nu-1)];
 *   - the padded work array W has (nni+8) x (nmu+8) entries with
 *     W(nu+4, mu+4) = f(nu, mu); N8 = nni+8 is the mu stride.
 *
 * Every function cites the reference routine it restates.  The arithmetic
 * follows the reference's expression order term by term; build with
 * -ffp-contract=off so that no fused multiply-adds are formed (the
 * reference build on x86-64 forms none either).
 *
 * Pinned by tests/test_oracle.py against vectors written by the reference's
 * own routines (tests/golden/*.npz, produced by oracle/ref_harness.f90).
 */
#include <math.h>
#include <string.h>
#include "x2oracle.h"

#define PIDX(nu, mu, N8) ((size_t)((mu) + 3) * (size_t)(N8) + (size_t)((nu) + 3))
#define UIDX(nu, mu, nni) ((size_t)((mu) - 1) * (size_t)(nni) + (size_t)((nu) - 1))

/* putin, inout.F90:23-82: embed f into W and fill the 4-wide mirror halos
 * below mu=1 and beyond nu=1 / nu=nni (negated for odd symmetry).  The rows
 * above mu=nmu are not touched (the reference leaves them as they were). */
void x2o_putin(int nni, int nmu, int isym, const double *fun, double *work)
{
    const int N8 = nni + 8;
    const double s = (isym == 1) ? 1.0 : -1.0;
    for (int mu = 1; mu <= nmu; ++mu)
        for (int nu = 1; nu <= nni; ++nu)
            work[PIDX(nu, mu, N8)] = fun[UIDX(nu, mu, nni)];
    /* mu = 0..-3 from mu = 2..5 (inout.F90:53-55) */
    for (int i = 2; i <= 5; ++i)
        for (int nu = 1; nu <= nni; ++nu)
            work[PIDX(nu, 2 - i, N8)] = s * fun[UIDX(nu, i, nni)];
    /* nu = 0..-3 from nu = 2..5 (inout.F90:57-59) */
    for (int j = 2; j <= 5; ++j)
        for (int mu = 1; mu <= nmu; ++mu)
            work[PIDX(2 - j, mu, N8)] = s * fun[UIDX(j, mu, nni)];
    /* nu = nni+1..nni+4 from nu = nni-1..nni-4 (inout.F90:61-63) */
    for (int j = 1; j <= 4; ++j)
        for (int mu = 1; mu <= nmu; ++mu)
            work[PIDX(nni + j, mu, N8)] = s * fun[UIDX(nni - j, mu, nni)];
}

/* putout, inout.F90:165-183 */
void x2o_putout(int nni, int nmu, double *fun, const double *work)
{
    const int N8 = nni + 8;
    for (int mu = 1; mu <= nmu; ++mu)
        for (int nu = 1; nu <= nni; ++nu)
            fun[UIDX(nu, mu, nni)] = work[PIDX(nu, mu, N8)];
}

/* One point update of the 17-point cross stencil, sormcsor.F90:43-71. */
static inline void update_point(double *f, size_t p, size_t u, int N8, const double *lhs,
                                const double *rhs, const double *b, const double *d,
                                const double *c, double omega, double omega1)
{
    const double *dmu1 = c, *dmu2 = c + 4, *dni1 = c + 8, *dni2 = c + 12;
    const size_t n1 = (size_t)N8, n2 = 2 * n1, n3 = 3 * n1, n4 = 4 * n1;
    double ddmi2 = dmu2[0] * (f[p - n4] + f[p + n4]) + dmu2[1] * (f[p - n3] + f[p + n3]) +
                   dmu2[2] * (f[p - n2] + f[p + n2]) + dmu2[3] * (f[p - n1] + f[p + n1]);
    double ddmi1 = dmu1[0] * (f[p - n4] - f[p + n4]) + dmu1[1] * (f[p - n3] - f[p + n3]) +
                   dmu1[2] * (f[p - n2] - f[p + n2]) + dmu1[3] * (f[p - n1] - f[p + n1]);
    double ddni2 = dni2[0] * (f[p - 4] + f[p + 4]) + dni2[1] * (f[p - 3] + f[p + 3]) +
                   dni2[2] * (f[p - 2] + f[p + 2]) + dni2[3] * (f[p - 1] + f[p + 1]);
    double ddni1 = dni1[0] * (f[p - 4] - f[p + 4]) + dni1[1] * (f[p - 3] - f[p + 3]) +
                   dni1[2] * (f[p - 2] - f[p + 2]) + dni1[3] * (f[p - 1] - f[p + 1]);
    double cc = ddmi2 + b[u] * ddmi1 + ddni2 + d[u] * ddni1;
    f[p] = omega * (rhs[u] - cc) / lhs[u] + omega1 * f[p];
}

/* Boundary lines after each sweep, sormcsor.F90:94-129 with the index sets of
 * mesh (initCBAllocArrays.F90:2823-2842): mu=1 for nu=2..nni-1 first, then
 * nu=1 and nu=nni for mu=1..nmu-4.  Even functions are extrapolated with the
 * 5-point exeven weights, odd functions are zero there. */
void x2o_boundary(int nni, int nmu, int isym, double *f, const double *exeven)
{
    const int N8 = nni + 8;
    const size_t n1 = (size_t)N8;
    if (isym == 1) {
        for (int nu = 2; nu <= nni - 1; ++nu) {
            size_t p = PIDX(nu, 1, N8);
            f[p] = exeven[0] * f[p + n1] + exeven[1] * f[p + 2 * n1] + exeven[2] * f[p + 3 * n1] +
                   exeven[3] * f[p + 4 * n1] + exeven[4] * f[p + 5 * n1];
        }
        for (int mu = 1; mu <= nmu - 4; ++mu) {
            size_t p = PIDX(1, mu, N8);
            f[p] = exeven[0] * f[p + 1] + exeven[1] * f[p + 2] + exeven[2] * f[p + 3] +
                   exeven[3] * f[p + 4] + exeven[4] * f[p + 5];
        }
        for (int mu = 1; mu <= nmu - 4; ++mu) {
            size_t p = PIDX(nni, mu, N8);
            f[p] = exeven[0] * f[p - 1] + exeven[1] * f[p - 2] + exeven[2] * f[p - 3] +
                   exeven[3] * f[p - 4] + exeven[4] * f[p - 5];
        }
    } else {
        for (int nu = 2; nu <= nni - 1; ++nu) f[PIDX(nu, 1, N8)] = 0.0;
        for (int mu = 1; mu <= nmu - 4; ++mu) f[PIDX(1, mu, N8)] = 0.0;
        for (int mu = 1; mu <= nmu - 4; ++mu) f[PIDX(nni, mu, N8)] = 0.0;
    }
}

/* mcsor, sormcsor.F90:143-386, with the col-wise mesh ordering
 * (initCBAllocArrays.F90:2467-2497): the relaxed points are enumerated with
 * mu = 2..nmu-4 outer and nu = 2..nni-1 inner, k = 0,1,2,...; pass c (c=0..4)
 * updates the points with k mod 5 == c in increasing k.  When nni is 1 mod 5
 * (every grid the reference admits) the pass index equals (nu - mu) mod 5.
 *
 * coef = {dmu1[4], dmu2[4], dni1[4], dni2[4]} (initSuppl,
 * initCBAllocArrays.F90:1979-1990).  mode 0: mcsor (five passes per sweep);
 * mode 1: sor (sormcsor.F90:20-132, one pass in col-wise order). */
void x2o_mcsor(int nni, int nmu, int isym, double *work, const double *lhs, const double *rhs,
               const double *b, const double *d, const double *coef, const double *exeven,
               double omega, int nsweeps, int mode)
{
    const int N8 = nni + 8;
    const double omega1 = 1.0 - omega;
    const int nrow = nni - 2;
    const long ngrid1 = (long)nrow * (long)(nmu - 5);
    for (int it = 0; it < nsweeps; ++it) {
        if (mode == 0) {
            for (int c = 0; c < 5; ++c)
                for (long k = c; k < ngrid1; k += 5) {
                    int mu = 2 + (int)(k / nrow), nu = 2 + (int)(k % nrow);
                    update_point(work, PIDX(nu, mu, N8), UIDX(nu, mu, nni), N8, lhs, rhs, b, d,
                                 coef, omega, omega1);
                }
        } else {
            for (long k = 0; k < ngrid1; ++k) {
                int mu = 2 + (int)(k / nrow), nu = 2 + (int)(k % nrow);
                update_point(work, PIDX(nu, mu, N8), UIDX(nu, mu, nni), N8, lhs, rhs, b, d, coef,
                             omega, omega1);
            }
        }
        x2o_boundary(nni, nmu, isym, work, exeven);
    }
}

/* Single-precision twin of x2o_mcsor (mode 0 only): the oracle for the FP32
 * correction sweeps.  Coefficients are rounded to float once, as the device
 * path does. */
void x2o_mcsor_f32(int nni, int nmu, int isym, float *f, const float *lhs, const float *rhs,
                   const float *b, const float *d, const double *coef, const double *exeven,
                   double omega, int nsweeps)
{
    const int N8 = nni + 8;
    const size_t n1 = (size_t)N8, n2 = 2 * n1, n3 = 3 * n1, n4 = 4 * n1;
    float c[16], ex[5];
    for (int i = 0; i < 16; ++i) c[i] = (float)coef[i];
    for (int i = 0; i < 5; ++i) ex[i] = (float)exeven[i];
    const float *dmu1 = c, *dmu2 = c + 4, *dni1 = c + 8, *dni2 = c + 12;
    const float om = (float)omega, om1 = (float)(1.0 - omega);
    const int nrow = nni - 2;
    const long ngrid1 = (long)nrow * (long)(nmu - 5);
    for (int it = 0; it < nsweeps; ++it) {
        for (int cc5 = 0; cc5 < 5; ++cc5)
            for (long k = cc5; k < ngrid1; k += 5) {
                int mu = 2 + (int)(k / nrow), nu = 2 + (int)(k % nrow);
                size_t p = PIDX(nu, mu, N8), u = UIDX(nu, mu, nni);
                float ddmi2 = dmu2[0] * (f[p - n4] + f[p + n4]) + dmu2[1] * (f[p - n3] + f[p + n3]) +
                              dmu2[2] * (f[p - n2] + f[p + n2]) + dmu2[3] * (f[p - n1] + f[p + n1]);
                float ddmi1 = dmu1[0] * (f[p - n4] - f[p + n4]) + dmu1[1] * (f[p - n3] - f[p + n3]) +
                              dmu1[2] * (f[p - n2] - f[p + n2]) + dmu1[3] * (f[p - n1] - f[p + n1]);
                float ddni2 = dni2[0] * (f[p - 4] + f[p + 4]) + dni2[1] * (f[p - 3] + f[p + 3]) +
                              dni2[2] * (f[p - 2] + f[p + 2]) + dni2[3] * (f[p - 1] + f[p + 1]);
                float ddni1 = dni1[0] * (f[p - 4] - f[p + 4]) + dni1[1] * (f[p - 3] - f[p + 3]) +
                              dni1[2] * (f[p - 2] - f[p + 2]) + dni1[3] * (f[p - 1] - f[p + 1]);
                float cv = ddmi2 + b[u] * ddmi1 + ddni2 + d[u] * ddni1;
                f[p] = om * (rhs[u] - cv) / lhs[u] + om1 * f[p];
            }
        if (isym == 1) {
            for (int nu = 2; nu <= nni - 1; ++nu) {
                size_t p = PIDX(nu, 1, N8);
                f[p] = ex[0] * f[p + n1] + ex[1] * f[p + n2] + ex[2] * f[p + n3] + ex[3] * f[p + n4] +
                       ex[4] * f[p + 5 * n1];
            }
            for (int mu = 1; mu <= nmu - 4; ++mu) {
                size_t p = PIDX(1, mu, N8);
                f[p] = ex[0] * f[p + 1] + ex[1] * f[p + 2] + ex[2] * f[p + 3] + ex[3] * f[p + 4] +
                       ex[4] * f[p + 5];
            }
            for (int mu = 1; mu <= nmu - 4; ++mu) {
                size_t p = PIDX(nni, mu, N8);
                f[p] = ex[0] * f[p - 1] + ex[1] * f[p - 2] + ex[2] * f[p - 3] + ex[3] * f[p - 4] +
                       ex[4] * f[p - 5];
            }
        } else {
            for (int nu = 2; nu <= nni - 1; ++nu) f[PIDX(nu, 1, N8)] = 0.0f;
            for (int mu = 1; mu <= nmu - 4; ++mu) f[PIDX(1, mu, N8)] = 0.0f;
            for (int mu = 1; mu <= nmu - 4; ++mu) f[PIDX(nni, mu, N8)] = 0.0f;
        }
    }
}

/* Residual r = rhs - (L + lhs) u on the relaxed interior (nu=2..nni-1,
 * mu=2..nmu-4), zero elsewhere: the defect the mixed-precision path forms in
 * FP64.  L is the same 17-point operator as in sormcsor.F90:46-70 and lhs
 * carries its diagonal. */
void x2o_residual(int nni, int nmu, const double *work, const double *lhs, const double *rhs,
                  const double *b, const double *d, const double *coef, double *res)
{
    const int N8 = nni + 8;
    const double *dmu1 = coef, *dmu2 = coef + 4, *dni1 = coef + 8, *dni2 = coef + 12;
    const size_t n1 = (size_t)N8, n2 = 2 * n1, n3 = 3 * n1, n4 = 4 * n1;
    memset(res, 0, sizeof(double) * (size_t)nni * (size_t)nmu);
    for (int mu = 2; mu <= nmu - 4; ++mu)
        for (int nu = 2; nu <= nni - 1; ++nu) {
            size_t p = PIDX(nu, mu, N8), u = UIDX(nu, mu, nni);
            const double *f = work;
            double ddmi2 = dmu2[0] * (f[p - n4] + f[p + n4]) + dmu2[1] * (f[p - n3] + f[p + n3]) +
                           dmu2[2] * (f[p - n2] + f[p + n2]) + dmu2[3] * (f[p - n1] + f[p + n1]);
            double ddmi1 = dmu1[0] * (f[p - n4] - f[p + n4]) + dmu1[1] * (f[p - n3] - f[p + n3]) +
                           dmu1[2] * (f[p - n2] - f[p + n2]) + dmu1[3] * (f[p - n1] - f[p + n1]);
            double ddni2 = dni2[0] * (f[p - 4] + f[p + 4]) + dni2[1] * (f[p - 3] + f[p + 3]) +
                           dni2[2] * (f[p - 2] + f[p + 2]) + dni2[3] * (f[p - 1] + f[p + 1]);
            double ddni1 = dni1[0] * (f[p - 4] - f[p + 4]) + dni1[1] * (f[p - 3] - f[p + 3]) +
                           dni1[2] * (f[p - 2] - f[p + 2]) + dni1[3] * (f[p - 1] - f[p + 1]);
            double cc = ddmi2 + b[u] * ddmi1 + ddni2 + d[u] * ddni1;
            res[u] = rhs[u] - (cc + lhs[u] * f[p]);
        }
}

/* orbAsymptGet, relaxOrbs.F90:415-437: decay factors for mu=nmu-3..nmu from
 * fa (nni x nmu; only rows nmu-4..nmu are read). edecay is nni x 4. */
void x2o_orb_asympt_get(int nni, int nmu, const double *vxisq, const double *vetasq,
                        const double *fa, double *edecay)
{
    for (int jj = 0; jj < 4; ++jj) {
        int j = nmu - 3 + jj;
        for (int i = 1; i <= nni; ++i) {
            double raiq1 = sqrt(vxisq[j - 2] + vetasq[i - 1] - 1.0);
            double raiq = sqrt(vxisq[j - 1] + vetasq[i - 1] - 1.0);
            double a = fa[UIDX(i, j, nni)], a1 = fa[UIDX(i, j - 1, nni)];
            edecay[(size_t)jj * nni + (i - 1)] =
                pow(fabs(a / a1), 0.25) * exp(sqrt(fabs(a1)) * (raiq1 - raiq));
        }
    }
}

/* orbAsymptSet, relaxOrbs.F90:444-462 */
void x2o_orb_asympt_set(int nni, int nmu, double *psi, const double *edecay)
{
    for (int jj = 0; jj < 4; ++jj) {
        int j = nmu - 3 + jj;
        for (int i = 1; i <= nni; ++i)
            psi[UIDX(i, j, nni)] = psi[UIDX(i, j - 1, nni)] * edecay[(size_t)jj * nni + (i - 1)];
    }
}

/* Sequence of relaxed points of mesh (initCBAllocArrays.F90:2408-2810) as
 * unpadded 0-based indices ik = (mu-1)*nni + (nu-1), for the reference's
 * orderings 0 col-wise (2467-2497: mu = 2..nmu-4 outer, nu = 2..nni-1
 * inner), 1 middle (2499-2540: rows from the middle row down to mu = 2, then
 * from the one above it up to nmu-4; in every row nu likewise from the middle
 * column), 2 middlemt (2543-2575: the middle split one row and column
 * further, the upper rows walked from nmu-4 back to the middle), 3 row-wise
 * (2578-2599: nu outer), 4 rrow-wis (2601-2627: row-wise reversed), 5 test1,
 * 7 test3 and 9 middle2 (2630-2665, 2698-2733, 2774-2810: the middle walks
 * with nu outer), 6 test2 (2667-2696: nu outer from the middle column, mu
 * from nmu-4 down to 2), 8 test4 (2735-2771: mu from nmu-4 down, nu from the
 * middle column).  Returns the count, (nni-2)*(nmu-5), or -1 for an unknown
 * ordering. */
long x2o_mesh(int nni, int nmu, int ordering, int *seq)
{
    /* padded rows i = 6..nmu (mu = i-4 = 2..nmu-4), columns j = 6..nni+3
     * (nu = 2..nni-1); the "mirrored" walks start at the middle row / column
     * and go outwards, first towards index 6, then towards the end
     * (initCBAllocArrays.F90:2458-2810) */
    const int nni8 = nni + 8, nmu8 = nmu + 8, iend1 = nmu, jend = nni + 3;
    const int nmux = (nmu8 - 1) / 2, nnix = (nni8 - 1) / 2;
    long k = 0;
#define EMIT(ii, jj) (seq[k++] = ((ii) - 5) * nni + ((jj) - 5))
#define MIR_I(i) ((i) <= nmux ? nmux - (i) + 6 : (i))
#define MIR_J(j) ((j) <= nnix ? nnix - (j) + 6 : (j))
    switch (ordering) {
    case 0: /* col-wise */
        for (int i = 6; i <= iend1; ++i)
            for (int j = 6; j <= jend; ++j) EMIT(i, j);
        break;
    case 1: /* middle */
        for (int i = 6; i <= iend1; ++i)
            for (int j = 6; j <= jend; ++j) EMIT(MIR_I(i), MIR_J(j));
        break;
    case 2: { /* middlemt: the upper half walks from the last row back to the middle */
        const int mx = nmux + 1, nx = nnix + 1;
        for (int i = 6; i <= iend1; ++i) {
            const int ii = i <= mx ? mx - i + 6 : nmu8 - 7 - (i - mx);
            for (int j = 6; j <= jend; ++j) EMIT(ii, j <= nx ? nx - j + 6 : j);
        }
        break;
    }
    case 3: /* row-wise */
        for (int j = 6; j <= jend; ++j)
            for (int i = 6; i <= iend1; ++i) EMIT(i, j);
        break;
    case 4: /* rrow-wis */
        for (int j = jend; j >= 6; --j)
            for (int i = iend1; i >= 6; --i) EMIT(i, j);
        break;
    case 5: /* test1 */
    case 7: /* test3 */
    case 9: /* middle2 */
        for (int j = 6; j <= jend; ++j)
            for (int i = 6; i <= iend1; ++i) EMIT(MIR_I(i), MIR_J(j));
        break;
    case 6: /* test2 */
        for (int j = 6; j <= jend; ++j)
            for (int i = iend1; i >= 6; --i) EMIT(i, MIR_J(j));
        break;
    case 8: /* test4 */
        for (int i = iend1; i >= 6; --i)
            for (int j = 6; j <= jend; ++j) EMIT(i, MIR_J(j));
        break;
    default:
        return -1;
    }
#undef EMIT
#undef MIR_I
#undef MIR_J
    return k;
}

/* sor (sormcsor.F90:20-132, mode 1) or mcsor (sormcsor.F90:143-386, mode 0:
 * five passes, pass c over the sequence positions i = c, c+5, c+10, ...)
 * of nsweeps sweeps over the points of mesh ordering `ordering` (x2o_mesh),
 * each sweep followed by the boundary lines.  Mode 2 is sor with the
 * reference's backward sweep (isstart = ngrid1, isstop = 1, isstep = -1 of
 * the altsweeps option, doSCF.F90:403-411, used by sormcsor.F90:42): the
 * sequence from its last position to its first. */
void x2o_relax_ordered(int nni, int nmu, int isym, double *work, const double *lhs, const double *rhs,
                       const double *b, const double *d, const double *coef, const double *exeven,
                       double omega, int nsweeps, int mode, int ordering, int *seq)
{
    const int N8 = nni + 8;
    const double omega1 = 1.0 - omega;
    const long n = x2o_mesh(nni, nmu, ordering, seq);
    for (int it = 0; it < nsweeps; ++it) {
        for (int c = 0; c < (mode == 0 ? 5 : 1); ++c)
            for (long k = c; k < n; k += (mode == 0 ? 5 : 1)) {
                const int ik = mode == 2 ? seq[n - 1 - k] : seq[k];
                const int mu = ik / nni + 1, nu = ik % nni + 1;
                update_point(work, PIDX(nu, mu, N8), (size_t)ik, N8, lhs, rhs, b, d, coef, omega, omega1);
            }
        x2o_boundary(nni, nmu, isym, work, exeven);
    }
}
