// Elementwise and quadrature kernels around the relaxation: the Fock
// potential assembly (fock.F90:57-184), the Newton-Cotes quadratures of
// norm/ortho/EaHF/etotalHF (ddot against wgt1/wgt2), the one-electron integral
// with the 9-point (nu,mu) Laplacian (integrals.F90:406-480,
// utils.f90:599-695), the multipole moments (scfUtils.F90:490-850), the
// multipole tails of the potentials (scfUtils.F90:275-329, 1039-1270) and the
// g/u symmetrisation (scfUtils.F90:380-406).
//
// All reductions are FP64, two-pass and deterministic: each workgroup reduces
// a fixed slice (wavefront shuffles, then LDS across the 4 wavefronts), and a
// second kernel adds the per-workgroup partials in a fixed order.
#include "x2gpu_internal.h"
#include <algorithm>
#include <cstring>

namespace x2g {

static inline int gx(long work, int block) { return (int)std::min<long>((work + block - 1) / block, 1024); }

template <typename D>
static int stage(x2g_ctx *ctx, const D *v, size_t cnt, size_t & /*unused*/, const D **out)
{
    return stage_n(ctx, v, cnt, out);
}

constexpr int kRedBlock = 256;   // 4 wavefronts
constexpr int kRedBlocks = 240;  // partials per term (fixed: deterministic order)

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block-wide sum of up to NV values per thread; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v[k] = wave_sum(v[k]);
        if (lane == 0) sh[w * NV + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = sh[k] + sh[NV + k] + sh[2 * NV + k] + sh[3 * NV + k];
    }
}

// ---------------------------------------------------------------------------
__global__ void k_fill(double *dst, size_t n, double v)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = v;
}


// out = bscale*base + sum_k c_k a_k b_k cc_k over n points, for the block's
// share of a grid-stride loop with blocks of 256 threads.  Each thread takes
// kLcIlp points 256 apart per pass, so a term's loads for all of them are in
// flight together (lanes past n read point n-1 and store nothing).  Per point
// the operations and their order are those of the plain loop.
constexpr int kLcIlp = 4;
__device__ __forceinline__ void lincomb_at(double *out, const double *base, double bscale,
                                           const x2g_term *__restrict__ t, int nt, long n, long e0)
{
    long e[kLcIlp];
    double v[kLcIlp];
#pragma unroll
    for (int j = 0; j < kLcIlp; ++j) e[j] = min(e0 + 256L * j, n - 1);
    if (base) {
#pragma unroll
        for (int j = 0; j < kLcIlp; ++j) v[j] = bscale * gmem(base)[e[j]];
    } else {
#pragma unroll
        for (int j = 0; j < kLcIlp; ++j) v[j] = 0.0;
    }
    for (int k = 0; k < nt; ++k) {
        const double c = t[k].c;
        const double *a = t[k].a_d, *b = t[k].b_d, *cc = t[k].cc_d;
        double x[kLcIlp];
#pragma unroll
        for (int j = 0; j < kLcIlp; ++j) x[j] = c * gmem(a)[e[j]];
        if (b) {
#pragma unroll
            for (int j = 0; j < kLcIlp; ++j) x[j] *= gmem(b)[e[j]];
        }
        if (cc) {
#pragma unroll
            for (int j = 0; j < kLcIlp; ++j) x[j] *= gmem(cc)[e[j]];
        }
#pragma unroll
        for (int j = 0; j < kLcIlp; ++j) v[j] += x[j];
    }
#pragma unroll
    for (int j = 0; j < kLcIlp; ++j)
        if (e0 + 256L * j < n) gmem_w(out)[e0 + 256L * j] = v[j];
}

__device__ __forceinline__ void lincomb_points(double *out, const double *base, double bscale,
                                               const x2g_term *__restrict__ t, int nt, long n, long first, long stride)
{
    for (long e0 = first; e0 < n; e0 += stride) lincomb_at(out, base, bscale, t, nt, n, e0);
}

__global__ void __launch_bounds__(256) k_lincomb(double *out, const double *base, double bscale,
                                                 const x2g_term *__restrict__ t, int nt, long n)
{
    lincomb_points(out, base, bscale, t, nt, n, (long)blockIdx.x * 256 * kLcIlp + threadIdx.x,
                   (long)gridDim.x * 256 * kLcIlp);
}

// partial[k*kRedBlocks + blk] = sum over the block's slice of w*a*b*c
__global__ void __launch_bounds__(kRedBlock) k_wdot(const x2g_dot *__restrict__ t, long n, double *partial)
{
    __shared__ double sh[4];
    const x2g_dot &d = t[blockIdx.y];
    double s[1] = {0.0};
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        double x = d.a_d[e];
        if (d.b_d) x *= d.b_d[e];
        if (d.cc_d) x *= d.cc_d[e];
        s[0] += d.w_d[e] * x;
    }
    block_sum<1>(s, sh);
    if (threadIdx.x == 0) partial[(long)blockIdx.y * gridDim.x + blockIdx.x] = s[0];
}

// out[k] = sum of the nb partials of term k (fixed order, one block per term)
__global__ void __launch_bounds__(kRedBlock) k_finish(const double *partial, int nb, int nv, double *out)
{
    __shared__ double sh[4];
    double s[1] = {0.0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s[0] += partial[(long)blockIdx.x * nb + i];
    block_sum<1>(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
    (void)nv;
}

// ---------------------------------------------------------------------------
// One-electron integral.  W is psi1 embedded with its mirror halos (putin;
// the rows above mu=nmu are zero).  For every grid point (nu,mu):
//   T psi1 = sum_k dni(k,nu) W(nu+k,mu) + sum_k dmu(k,mu) W(nu,mu+k),
//   dni(k,nu) = a2n(k) + d(nu) a1n(k),  dmu(k,mu) = a2m(k) + borb(mu) a1m(k)
// (9-point first/second derivative weights, initSuppl
// initCBAllocArrays.F90:1993-2032), then
//   kin  += wgt1*psi2*(T psi1 + m^2 e psi1),   nuc += wgt1*psi2*f0*psi1.
struct OneelArgs {
    const double *w;
    const double *psi1, *psi2, *wgt1, *e, *f0, *bmu, *dnu;
    double a1n[9], a2n[9], a1m[9], a2m[9];
    double m2;
    int nni, nmu;
};

__global__ void __launch_bounds__(kRedBlock) k_oneel(OneelArgs A, double *partial)
{
    __shared__ double sh[8];
    const int nni = A.nni, nmu = A.nmu;
    const long s1 = nni + 8;
    const long n = (long)nni * nmu;
    double acc[2] = {0.0, 0.0};
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int mu = 1 + (int)(e / nni), nu = 1 + (int)(e % nni);
        const long p = (long)(mu + 3) * s1 + nu + 3;
        const double dn = A.dnu[nu - 1], bm = A.bmu[mu - 1];
        double tn = 0.0, tm = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            tn += A.w[p + (k - 4)] * (A.a2n[k] + dn * A.a1n[k]);
            tm += A.w[p + (long)(k - 4) * s1] * (A.a2m[k] + bm * A.a1m[k]);
        }
        double t = tn + tm;
        const double ps1 = A.psi1[e], ps2 = A.psi2[e], wg = A.wgt1[e];
        if (A.m2 != 0.0) t += A.m2 * A.e[e] * ps1;
        acc[0] += wg * (t * ps2);
        acc[1] += wg * (ps1 * A.f0[e] * ps2);
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = acc[0];
        partial[gridDim.x + blockIdx.x] = acc[1];
    }
}

// ---------------------------------------------------------------------------
// Normalised associated Legendre values N_kq P_k^q(x), k = 1..kmax, with the
// Condon-Shortley phase, N_kq = sqrt((k-q)!/(k+q)!); standard upward
// recurrence in k from P_q^q = (-1)^q (2q-1)!! (1-x^2)^(q/2).
// `quirk3` reproduces the reference tail formula for q = 3, which raises
// (1-x^2) to the power 2 instead of 3/2 (vexch, scfUtils.F90:1160-1163).
__device__ inline void assoc_legendre(int q, double x, int kmax, double *out, bool quirk3)
{
    for (int k = 0; k < 8; ++k) out[k] = 0.0;
    const double s2 = fabs(1.0 - x * x);
    double pmm = 1.0;  // P_q^q
    for (int i = 1; i <= q; ++i) pmm *= -(double)(2 * i - 1);
    if (q > 0) {
        double sp = (q == 3 && quirk3) ? s2 * s2 : pow(s2, 0.5 * q);
        pmm *= sp;
    }
    double pkm2 = 0.0, pkm1 = pmm;  // P_{q-1}^q = 0, P_q^q
    for (int k = q; k <= kmax; ++k) {
        double pk;
        if (k == q) pk = pmm;
        else pk = ((double)(2 * k - 1) * x * pkm1 - (double)(k + q - 1) * pkm2) / (double)(k - q);
        if (k > q) { pkm2 = pkm1; pkm1 = pk; }
        // normalisation sqrt((k-q)!/(k+q)!)
        double nrm = 1.0;
        for (int j = k - q + 1; j <= k + q; ++j) nrm /= (double)j;
        if (k >= 1) out[k - 1] = pk * sqrt(nrm);
    }
}

// Plain Legendre P_1..P_kmax in the reference's recurrence order
// (initLP, initCBAllocArrays.F90:1547-1552).
__device__ inline void legendre_ref(double c, int kmax, double *out)
{
    for (int k = 0; k < 8; ++k) out[k] = 0.0;
    out[0] = c;
    out[1] = (3.0 * c * c - 1.0) * 0.5;
    for (int n = 2; n <= kmax - 1; ++n) out[n] = ((double)(2 * n + 1) * c * out[n - 1] - (double)n * out[n - 2]) / (double)(n + 1);
}

struct MomArgs {
    const double *psi1, *psi2, *wgt2, *f4, *vxi, *vxisq, *veta, *vetasq;
    double r2;
    int nni, nmu, dm, mpole;
};

__device__ __forceinline__ void moments_body(const MomArgs &A, double *partial)
{
    __shared__ double sh[32];
    const int nni = A.nni;
    const long n = (long)nni * A.nmu;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const double eps = 2.220446049250313e-16;
    // 32-bit index arithmetic (a grid has fewer than 2^31 points)
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < (int)n; e += gridDim.x * blockDim.x) {
        const int mu = e / nni, nu = e - mu * nni;
        const double rr = sqrt(A.vxisq[mu] + A.vetasq[nu] - 1.0);
        double c = 0.0;
        if (A.dm == 0) { if (!(fabs(rr) < eps)) c = A.vxi[mu] * A.veta[nu] / rr; }
        else { if (fabs(rr) > eps) c = A.veta[nu] * A.vxi[mu] / rr; }
        double P[8];
        if (A.dm == 0) legendre_ref(c, A.mpole, P);
        else assoc_legendre(A.dm, c, A.mpole, P, false);
        const double xr = A.r2 * rr;
        const double dens = A.f4[e] * (A.psi1[e] * A.psi2[e]);
        const double wd = A.wgt2[e];
        double xw = 1.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            xw *= xr;
            if (k < A.mpole) acc[k] += wd * ((P[k] * xw) * dens);
        }
    }
    block_sum<8>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) partial[(long)k * gridDim.x + blockIdx.x] = acc[k];
}

__global__ void __launch_bounds__(kRedBlock) k_moments(MomArgs A, double *partial) { moments_body(A, partial); }

// blockIdx.y = pair; partial[(pair*8 + k)*gridDim.x + block]
__global__ void __launch_bounds__(kRedBlock) k_moments_batch(const MomArgs *__restrict__ A, double *partial)
{
    moments_body(A[blockIdx.y], partial + (long)blockIdx.y * 8 * gridDim.x);
}

// Potential tail rows mu = nmu-3..nmu.
struct TailArgs {
    double *pot;
    const double *vxi, *vxisq, *veta, *vetasq;
    double mom[8];
    double r2;
    int nni, nmu, kind, dm, mpole;
};

__device__ __forceinline__ void pot_tail_body(const TailArgs &A)
{
    const int nni = A.nni;
    const double eps = 2.220446049250313e-16;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < 4 * nni; e += gridDim.x * blockDim.x) {
        const int j = A.nmu - 3 + e / nni;  // 1-based mu
        const int i = 1 + e % nni;
        const double rr = sqrt(A.vxisq[j - 1] + A.vetasq[i - 1] - 1.0);
        double v;
        if (A.kind == 0) {
            // coulAsympt + vcoul (scfUtils.F90:275-296, 1039-1068)
            const double rr1 = 1.0 / (rr * A.r2);
            const double c = A.veta[i - 1] * A.vxi[j - 1] / rr;
            double P[8];
            legendre_ref(c, A.mpole, P);
            double pe = 0.0;
            for (int m = 1; m <= A.mpole; ++m) pe += A.mom[m - 1] * P[m - 1] * pow(rr1, (double)(m + 1));
            v = A.r2 * A.vxi[j - 1] * (rr1 + pe);
        } else {
            // exchAsympt + vexch (scfUtils.F90:310-329, 1075-1270)
            double c = 0.0;
            if (fabs(rr) > eps) c = A.veta[i - 1] * A.vxi[j - 1] / rr;
            const double rrs = rr * A.r2;
            double P[8];
            if (A.dm == 0) legendre_ref(c, A.mpole, P);
            else assoc_legendre(A.dm, c, A.mpole, P, true);
            double xrn = rrs * rrs, pe = 0.0;
            for (int m = 1; m <= A.mpole; ++m) {
                if (m > 1) xrn *= rrs;
                pe = pe + P[m - 1] * A.mom[m - 1] / xrn;
            }
            v = (A.dm > 4) ? 0.0 : A.r2 * A.vxi[j - 1] * pe;
        }
        A.pot[(long)(j - 1) * nni + i - 1] = v;
    }
}

__global__ void k_pot_tail(TailArgs A) { pot_tail_body(A); }
__global__ void k_pot_tail_batch(const TailArgs *__restrict__ A) { pot_tail_body(A[blockIdx.y]); }

__global__ void k_setci(double *psi, int nni, int nmu, int ihsym)
{
    const int n2 = nni / 2 + 1;
    const long tot = (long)(n2 - 1) * nmu;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int mi = (int)(e / (n2 - 1)), ni = 1 + (int)(e % (n2 - 1));
        const double v = psi[(long)mi * nni + ni - 1];
        psi[(long)mi * nni + (nni - ni)] = ihsym == 1 ? v : -v;
        if (ihsym != 1 && ni == 1) psi[(long)mi * nni + n2 - 1] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Batched kernels: blockIdx.y selects the problem.  The reductions keep the
// slicing of the single-problem kernels (kRedBlocks grid-strided slices, then
// a fixed-order sum), so a batched quadrature equals the single one bit for bit.

struct LcOp {
    double *out;
    const double *base;
    double bscale;
    int first, nterms;
    long n;
};

// One workgroup per (op, chunk of 256*kLcIlp points).  The 8 XCDs take a
// launch's workgroups round-robin, each with its own L2: XCD x gets chunks
// [x*slice, (x+1)*slice) and walks them chunk by chunk with the op index
// fastest, so the ops that share input arrays (a job's Coulomb and exchange
// potentials and orbitals) read a chunk's segment of each from that XCD's L2
// after the first of them.
__global__ void __launch_bounds__(256) k_lincomb_batch(const LcOp *__restrict__ ops, const x2g_term *__restrict__ terms,
                                                       int nops, int nchunk, int slice)
{
    const int b = blockIdx.x, x = b & 7, j = b >> 3;
    const int chunk = x * slice + j / nops, q = j % nops;
    if (chunk >= nchunk) return;
    const LcOp &o = ops[q];
    const long e0 = (long)chunk * 256 * kLcIlp + threadIdx.x;
    if ((long)chunk * 256 * kLcIlp >= o.n) return;
    lincomb_at(o.out, o.base, o.bscale, terms + o.first, o.nterms, o.n, e0);
}

// Sum of the nb partials at p (fixed order), valid in thread 0.
__device__ __forceinline__ double sum_partials(const double *p, int nb, double *sh)
{
    double s[1] = {0.0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s[0] += p[i];
    block_sum<1>(s, sh);
    __syncthreads();  // sh is reused by the caller's next sum
    return s[0];
}

constexpr int kMaxOrtho = X2G_MAX_ORTHO;
struct NoDesc {
    double *psi;
    const double *w, *f;
    const double *oth[kMaxOrtho];
    double *norm_out;
    long n;
    int nni, nmu, noth, ihomo, normalise;
};
// partial layout: ((item*(kMaxOrtho+1) + k)*kRedBlocks + block), k = 0 norm,
// k = 1+q overlap with other q

// The g/u symmetry (setCiOrb, scfUtils.F90:380-406) as a view: for ihomo != 0
// the nu > n2 = (nni+1)/2 half of a row is the mirror image of the nu < n2
// half (times ihomo), and for ihomo = -1 the middle column is zero.  Grids
// have odd nni.  psi_sym(e) reads only the nu <= n2 half.
__device__ __forceinline__ double psi_sym(const NoDesc &d, long e)
{
    const double *psi = d.psi;
    if (d.ihomo == 0) return gmem(psi)[e];
    const int nni = d.nni, n2 = nni / 2 + 1;
    const int mi = (int)(e / nni), nu = 1 + (int)(e - (long)mi * nni);
    if (nu < n2) return gmem(psi)[e];
    if (nu == n2) return d.ihomo == 1 ? gmem(psi)[e] : 0.0;
    const double v = gmem(psi)[(long)mi * nni + (nni - nu)];  // nu' = nni + 1 - nu
    return d.ihomo == 1 ? v : -v;
}

// blockIdx.z = 0: s = sum w*(psi*psi*f) over the orbital as it arrives (the
// reference normalises before imposing the symmetry); blockIdx.z = 1+q:
// sum w*(psi_sym*oth_q*f) with the symmetry imposed.
// Terms k0 .. k0+kNoGroup-1 of one item per workgroup (blockIdx.z = group):
// term 0 is <psi|psi>, term k >= 1 the overlap with other k-1.  The terms of
// a group share one pass over psi, f and w; each term's partial sums the
// same elements in the same order as a workgroup of its own would.
constexpr int kNoGroup = 4;
__global__ void __launch_bounds__(kRedBlock) k_no_dot(const NoDesc *__restrict__ nd, double *partial)
{
    __shared__ double sh[4 * kNoGroup];
    const NoDesc &d = nd[blockIdx.y];
    if (!d.normalise) return;
    const int k0 = blockIdx.z * kNoGroup;
    if (k0 > d.noth) return;
    const int nk = min(kNoGroup, d.noth + 1 - k0);
    const bool self = k0 == 0, overlaps = k0 + nk > 1;
    const double *oth[kNoGroup];
#pragma unroll
    for (int g = 0; g < kNoGroup; ++g) oth[g] = g < nk && k0 + g >= 1 ? d.oth[k0 + g - 1] : nullptr;
    double s[kNoGroup] = {0.0, 0.0, 0.0, 0.0};
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (long)gridDim.x * blockDim.x) {
        const double f = gmem(d.f)[e], w = gmem(d.w)[e];
        const double p = self ? gmem(d.psi)[e] : 0.0;
        const double ps = overlaps ? psi_sym(d, e) : 0.0;
#pragma unroll
        for (int g = 0; g < kNoGroup; ++g) {
            if (g >= nk) break;
            double x;
            if (self && g == 0) {
                x = p;
                x *= p;
            } else {
                x = ps;
                x *= gmem(oth[g])[e];
            }
            x *= f;
            s[g] += w * x;
        }
    }
    block_sum<kNoGroup>(s, sh);
    if (threadIdx.x == 0)
        for (int g = 0; g < nk; ++g)
            partial[((long)blockIdx.y * (kMaxOrtho + 1) + k0 + g) * gridDim.x + blockIdx.x] = s[g];
}

// The update psi <- c0 psi_sym + sum_k c_k other_k (with normalise = 0: psi
// <- psi_sym).  Each workgroup forms the coefficients itself from the dot
// partials (the same fixed-order sums in every workgroup, so all agree bit
// for bit); workgroup 0 writes the norm.  With the symmetry, a thread owns a
// point nu <= n2 and its mirror image, so the half it reads is written only
// by itself.
__global__ void __launch_bounds__(kRedBlock) k_no_apply(const NoDesc *__restrict__ nd, const double *partial, int nb)
{
    __shared__ double sh[4];
    __shared__ double c[kMaxOrtho + 1];
    const NoDesc &d = nd[blockIdx.y];
    if (!d.normalise && d.ihomo == 0) return;
    const int no = d.normalise ? d.noth : 0;
    if (d.normalise) {
        const double *p = partial + (long)blockIdx.y * (kMaxOrtho + 1) * nb;
        for (int k = 0; k <= no; ++k) {
            const double v = sum_partials(p + (long)k * nb, nb, sh);
            if (threadIdx.x == 0) c[k] = v;
        }
        if (threadIdx.x == 0) {
            const double s = c[0];
            if (d.norm_out && blockIdx.x == 0) *d.norm_out = s;
            const double rs = 1.0 / sqrt(s);
            double a = 0.0;
            for (int k = 1; k <= no; ++k) {
                c[k] = c[k] * rs;
                a += c[k] * c[k];
            }
            const double ano = no > 0 ? sqrt(1.0 + a) / (1.0 - a * a) : 1.0;
            c[0] = ano * rs;
            for (int k = 1; k <= no; ++k) c[k] = -ano * c[k];
        }
    } else if (threadIdx.x == 0) {
        c[0] = 1.0;
    }
    __syncthreads();
    const double *psi = d.psi;
    if (d.ihomo == 0) {
        for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (long)gridDim.x * blockDim.x) {
            double v = c[0] * gmem(psi)[e];
            for (int k = 0; k < no; ++k) v += c[1 + k] * gmem(d.oth[k])[e];
            gmem_w(d.psi)[e] = v;
        }
        return;
    }
    const int nni = d.nni, n2 = nni / 2 + 1;
    const long tot = (long)n2 * d.nmu;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const int mi = (int)(t / n2), nu = 1 + (int)(t - (long)mi * n2);
        const long e = (long)mi * nni + nu - 1;
        const double v1 = gmem(psi)[e];
        if (nu < n2) {
            const long e2 = (long)mi * nni + (nni - nu);  // nu' = nni + 1 - nu
            double a = c[0] * v1, b = c[0] * (d.ihomo == 1 ? v1 : -v1);
            for (int k = 0; k < no; ++k) {
                a += c[1 + k] * gmem(d.oth[k])[e];
                b += c[1 + k] * gmem(d.oth[k])[e2];
            }
            gmem_w(d.psi)[e] = a;
            gmem_w(d.psi)[e2] = b;
        } else {
            double a = c[0] * (d.ihomo == 1 ? v1 : 0.0);
            for (int k = 0; k < no; ++k) a += c[1 + k] * gmem(d.oth[k])[e];
            gmem_w(d.psi)[e] = a;
        }
    }
}

// EaHF of one orbital: the accumulators of k_oneel (kinetic, nuclear) and of
// the two x2g_wdot terms <psi|cpot psi> and <psi|xpot> in one pass.
struct EaDesc {
    OneelArgs a;
    const double *cpot, *xpot, *wgt2;
    double *ee_out;
};

__global__ void __launch_bounds__(kRedBlock) k_eahf(const EaDesc *__restrict__ ed, double *partial)
{
    __shared__ double sh[16];
    const EaDesc &D = ed[blockIdx.y];
    const OneelArgs &A = D.a;
    const int nni = A.nni, nmu = A.nmu;
    const long s1 = nni + 8;
    const long n = (long)nni * nmu;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // 32-bit index arithmetic (a grid has fewer than 2^31 points)
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < (int)n; e += gridDim.x * blockDim.x) {
        const int mu0 = e / nni, mu = 1 + mu0, nu = 1 + (e - mu0 * nni);
        const int p = (mu + 3) * (int)s1 + nu + 3;
        const double dn = A.dnu[nu - 1], bm = A.bmu[mu - 1];
        double tn = 0.0, tm = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            tn += A.w[p + (k - 4)] * (A.a2n[k] + dn * A.a1n[k]);
            tm += A.w[p + (long)(k - 4) * s1] * (A.a2m[k] + bm * A.a1m[k]);
        }
        double t = tn + tm;
        const double ps1 = A.psi1[e], ps2 = A.psi2[e], wg = A.wgt1[e];
        if (A.m2 != 0.0) t += A.m2 * A.e[e] * ps1;
        acc[0] += wg * (t * ps2);
        acc[1] += wg * (ps1 * A.f0[e] * ps2);
        double x = ps1 * D.cpot[e];
        x *= ps1;
        acc[2] += D.wgt2[e] * x;
        acc[3] += D.wgt2[e] * (ps1 * D.xpot[e]);
    }
    block_sum<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) partial[((long)blockIdx.y * 4 + k) * gridDim.x + blockIdx.x] = acc[k];
}

__global__ void __launch_bounds__(kRedBlock) k_eahf_finish(const EaDesc *__restrict__ ed, const double *partial, int nb)
{
    __shared__ double sh[4];
    __shared__ double r[4];
    const double *p = partial + (long)blockIdx.x * 4 * nb;
    for (int k = 0; k < 4; ++k) {
        const double v = sum_partials(p + (long)k * nb, nb, sh);
        if (threadIdx.x == 0) r[k] = v;
    }
    if (threadIdx.x == 0) *ed[blockIdx.x].ee_out = (r[0] + r[1]) + (r[2] - r[3]);
}

struct GatherDesc {
    const double *src;
    int count, offset;
};
__global__ void k_gather(const GatherDesc *__restrict__ gd, double *dst)
{
    const GatherDesc &g = gd[blockIdx.x];
    for (int i = threadIdx.x; i < g.count; i += blockDim.x) dst[g.offset + i] = g.src[i];
}

// 9-point weights (initSuppl, initCBAllocArrays.F90:1670-1673, 1993-2011):
// a1 = {3,-32,168,-672,0,672,-168,32,-3}/(840 h), a2 = {-9,128,-1008,8064,
// -14350,8064,-1008,128,-9}/(5040 h^2).
static void deriv_weights(double h, double *a1, double *a2)
{
    static const double aa1[9] = {3.0, -32.0, 168.0, -672.0, 0.0, 672.0, -168.0, 32.0, -3.0};
    static const double aa2[9] = {-9.0, 128.0, -1008.0, 8064.0, -14350.0, 8064.0, -1008.0, 128.0, -9.0};
    const double w1 = 1.0 / 840.0, w2 = 1.0 / 5040.0;
    for (int k = 0; k < 9; ++k) {
        a2[k] = w2 * aa2[k] / (h * h);
        a1[k] = w1 * aa1[k] / h;
    }
}

}  // namespace x2g

using namespace x2g;

extern "C" int x2g_fill(x2g_ctx *ctx, double *dst_d, size_t n, double value)
{
    X2G_CHECK(ctx && (dst_d || n == 0), "x2g_fill: null argument");
    if (n == 0) return X2G_OK;
    X2G_LAUNCH(ctx, k_fill, dim3(gx((long)n, 256)), dim3(256), 0, dst_d, n, value);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_lincomb(x2g_grid *g, double *out_d, const double *base_d, double base_scale,
                           const x2g_term *terms, int nterms)
{
    X2G_CHECK(g && out_d && nterms >= 0 && (terms || nterms == 0), "x2g_lincomb: null argument");
    for (int k = 0; k < nterms; ++k) X2G_CHECK(terms[k].a_d, "x2g_lincomb: term %d has no a", k);
    x2g_ctx *ctx = g->ctx;
    size_t off = 0;
    const x2g_term *dt = nullptr;
    if (nterms) {
        int rc = stage(ctx, terms, (size_t)nterms, off, &dt);
        if (rc) return rc;
    }
    X2G_LAUNCH(ctx, k_lincomb, dim3(gx((long)g->n(), 256 * kLcIlp)), dim3(256), 0, out_d, base_d, base_scale,
                       dt, nterms, (long)g->n());
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

static int reduce_finish(x2g_ctx *ctx, double *partial, int nv, double *out_h)
{
    double *dout = partial + (size_t)nv * kRedBlocks;
    X2G_LAUNCH(ctx, k_finish, dim3(nv), dim3(kRedBlock), 0, partial, kRedBlocks, nv, dout);
    X2G_HIP(hipGetLastError());
    X2G_HIP(hipMemcpyAsync(out_h, dout, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    return X2G_OK;
}

extern "C" int x2g_wdot(x2g_grid *g, const x2g_dot *terms, int nterms, double *out_h)
{
    X2G_CHECK(g && terms && out_h && nterms > 0, "x2g_wdot: null argument");
    for (int k = 0; k < nterms; ++k) X2G_CHECK(terms[k].w_d && terms[k].a_d, "x2g_wdot: term %d incomplete", k);
    x2g_ctx *ctx = g->ctx;
    int rc = ws_reserve(ctx, 1, (size_t)nterms * (kRedBlocks + 1) * sizeof(double));
    if (rc) return rc;
    size_t off = 0;
    const x2g_dot *dt;
    if ((rc = stage(ctx, terms, (size_t)nterms, off, &dt))) return rc;
    double *partial = (double *)ctx->ws[1].ptr;
    X2G_LAUNCH(ctx, k_wdot, dim3(kRedBlocks, nterms), dim3(kRedBlock), 0, dt, (long)g->n(), partial);
    X2G_HIP(hipGetLastError());
    return reduce_finish(ctx, partial, nterms, out_h);
}

extern "C" int x2g_oneel(x2g_grid *g, const double *psi1_d, const double *psi2_d, int isym, int m, double *out_h)
{
    X2G_CHECK(g && psi1_d && psi2_d && out_h, "x2g_oneel: null argument");
    X2G_CHECK(isym == 1 || isym == -1, "x2g_oneel: isym");
    x2g_ctx *ctx = g->ctx;
    const size_t pb = (g->npad() * sizeof(double) + 255) & ~size_t(255);
    int rc = ws_reserve(ctx, 0, pb);
    if (rc) return rc;
    if ((rc = ws_reserve(ctx, 1, 3 * (kRedBlocks + 1) * sizeof(double)))) return rc;
    double *w = (double *)ctx->ws[0].ptr;
    if ((rc = launch_putin(g, isym, psi1_d, w, nullptr))) return rc;
    OneelArgs A{};
    A.w = w;
    A.psi1 = psi1_d;
    A.psi2 = psi2_d;
    A.wgt1 = g->suppl + 12 * g->n();
    A.e = g->suppl + 3 * g->n();
    A.f0 = g->suppl + 4 * g->n();
    A.bmu = g->bmu_orb;
    A.dnu = g->dnu;
    deriv_weights(g->coef[22], A.a1n, A.a2n);  // hni
    deriv_weights(g->coef[23], A.a1m, A.a2m);  // hmu
    A.m2 = (double)(m * m);
    A.nni = g->nni;
    A.nmu = g->nmu;
    double *partial = (double *)ctx->ws[1].ptr;
    X2G_LAUNCH(ctx, k_oneel, dim3(kRedBlocks), dim3(kRedBlock), 0, A, partial);
    X2G_HIP(hipGetLastError());
    double r[2];
    if ((rc = reduce_finish(ctx, partial, 2, r))) return rc;
    out_h[0] = r[0] + r[1];
    out_h[1] = r[0];
    out_h[2] = r[1];
    return X2G_OK;
}

extern "C" int x2g_moments(x2g_grid *g, const double *psi1_d, const double *psi2_d, int dm, int mpole, double *out_h)
{
    X2G_CHECK(g && psi1_d && psi2_d && out_h, "x2g_moments: null argument");
    X2G_CHECK(mpole >= 2 && mpole <= 8, "x2g_moments: mpole must be 2..8");
    X2G_CHECK(dm >= 0 && dm <= 4, "x2g_moments: |Delta m| must be 0..4");
    x2g_ctx *ctx = g->ctx;
    int rc = ws_reserve(ctx, 1, 8 * (kRedBlocks + 1) * sizeof(double));
    if (rc) return rc;
    MomArgs A{};
    A.psi1 = psi1_d;
    A.psi2 = psi2_d;
    A.wgt2 = g->suppl + 13 * g->n();
    A.f4 = g->suppl + 8 * g->n();
    A.vxi = g->vxi;
    A.vxisq = g->vxisq;
    A.veta = g->veta;
    A.vetasq = g->vetasq;
    A.r2 = 0.5 * g->coef[24];
    A.nni = g->nni;
    A.nmu = g->nmu;
    A.dm = dm;
    A.mpole = mpole;
    double *partial = (double *)ctx->ws[1].ptr;
    X2G_LAUNCH(ctx, k_moments, dim3(kRedBlocks), dim3(kRedBlock), 0, A, partial);
    X2G_HIP(hipGetLastError());
    double r[8];
    if ((rc = reduce_finish(ctx, partial, 8, r))) return rc;
    for (int k = 0; k < mpole; ++k) out_h[k] = r[k];
    return X2G_OK;
}

extern "C" int x2g_pot_tail(x2g_grid *g, double *pot_d, int kind, int dm, int mpole, const double *mom)
{
    X2G_CHECK(g && pot_d && mom, "x2g_pot_tail: null argument");
    X2G_CHECK(kind == 0 || kind == 1, "x2g_pot_tail: kind");
    X2G_CHECK(mpole >= 2 && mpole <= 8, "x2g_pot_tail: mpole must be 2..8");
    X2G_CHECK(dm >= 0, "x2g_pot_tail: dm < 0");
    TailArgs A{};
    A.pot = pot_d;
    A.vxi = g->vxi;
    A.vxisq = g->vxisq;
    A.veta = g->veta;
    A.vetasq = g->vetasq;
    for (int k = 0; k < 8; ++k) A.mom[k] = k < mpole ? mom[k] : 0.0;
    A.r2 = 0.5 * g->coef[24];
    A.nni = g->nni;
    A.nmu = g->nmu;
    A.kind = kind;
    A.dm = dm;
    A.mpole = mpole;
    X2G_LAUNCH(g->ctx, k_pot_tail, dim3(gx(4L * g->nni, 128)), dim3(128), 0, A);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_setci(x2g_grid *g, double *psi_d, int ihsym)
{
    X2G_CHECK(g && psi_d, "x2g_setci: null argument");
    X2G_CHECK(ihsym == 1 || ihsym == -1, "x2g_setci: ihsym");
    const long tot = (long)(g->nni / 2) * g->nmu;
    X2G_LAUNCH(g->ctx, k_setci, dim3(gx(tot, 256)), dim3(256), 0, psi_d, g->nni, g->nmu, ihsym);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

// ---------------------------------------------------------------------------
// Batched C ABI

static int check_grids(x2g_ctx *ctx, x2g_grid *const *grids, int gi)
{
    X2G_CHECK(gi >= 0 && grids[gi], "bad grid index %d", gi);
    X2G_CHECK(grids[gi]->ctx == ctx, "grid of another context");
    return X2G_OK;
}

extern "C" int x2g_lincomb_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_lincomb_op *ops, int nops,
                                 const x2g_term *terms, int nterms_total)
{
    X2G_CHECK(ctx && grids && (ops || nops == 0) && nops >= 0 && nterms_total >= 0, "x2g_lincomb_batch: null argument");
    if (nops == 0) return X2G_OK;
    std::vector<LcOp> v(nops);
    long maxn = 0;
    for (int i = 0; i < nops; ++i) {
        const x2g_lincomb_op &o = ops[i];
        int rc = check_grids(ctx, grids, o.grid_index);
        if (rc) return rc;
        X2G_CHECK(o.out_d && o.nterms >= 0 && o.first_term >= 0 && o.first_term + o.nterms <= nterms_total,
                  "x2g_lincomb_batch: op %d has a bad term range", i);
        for (int k = 0; k < o.nterms; ++k)
            X2G_CHECK(terms[o.first_term + k].a_d, "x2g_lincomb_batch: term %d has no a", o.first_term + k);
        v[i] = {o.out_d, o.base_d, o.base_scale, o.first_term, o.nterms, (long)grids[o.grid_index]->n()};
        maxn = std::max(maxn, v[i].n);
    }
    size_t off = 0;
    const LcOp *dops;
    const x2g_term *dt = nullptr;
    int rc = stage(ctx, v.data(), v.size(), off, &dops);
    if (rc) return rc;
    if (nterms_total && (rc = stage(ctx, terms, (size_t)nterms_total, off, &dt))) return rc;
    const int nchunk = (int)((maxn + 256 * kLcIlp - 1) / (256 * kLcIlp)), slice = (nchunk + 7) / 8;
    X2G_LAUNCH(ctx, k_lincomb_batch, dim3(8 * nops * slice), dim3(256), 0, dops, dt, nops, nchunk, slice);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_normortho_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_normortho_item *items, int n)
{
    X2G_CHECK(ctx && grids && (items || n == 0) && n >= 0, "x2g_normortho_batch: null argument");
    if (n == 0) return X2G_OK;
    std::vector<NoDesc> v(n);
    long maxn = 0;
    int maxoth = 0;
    for (int i = 0; i < n; ++i) {
        const x2g_normortho_item &it = items[i];
        int rc = check_grids(ctx, grids, it.grid_index);
        if (rc) return rc;
        x2g_grid *g = grids[it.grid_index];
        X2G_CHECK(it.psi_d, "x2g_normortho_batch: item %d has no orbital", i);
        X2G_CHECK(it.nothers >= 0 && it.nothers <= kMaxOrtho, "x2g_normortho_batch: item %d nothers", i);
        X2G_CHECK(it.ihomo >= -1 && it.ihomo <= 1, "x2g_normortho_batch: item %d ihomo", i);
        X2G_CHECK(it.ihomo == 0 || g->nni % 2 == 1, "x2g_normortho_batch: item %d: symmetry needs odd nni", i);
        NoDesc &d = v[i];
        d.psi = it.psi_d;
        d.w = g->suppl + 13 * g->n();  // wgt2
        d.f = g->suppl + 8 * g->n();   // f4
        for (int k = 0; k < kMaxOrtho; ++k) d.oth[k] = k < it.nothers ? it.others_d[k] : nullptr;
        for (int k = 0; k < it.nothers; ++k) X2G_CHECK(d.oth[k], "x2g_normortho_batch: item %d other %d", i, k);
        d.norm_out = it.norm_out_d;
        d.n = (long)g->n();
        d.nni = g->nni;
        d.nmu = g->nmu;
        d.noth = it.normalise ? it.nothers : 0;
        d.ihomo = it.ihomo;
        d.normalise = it.normalise ? 1 : 0;
        maxn = std::max(maxn, d.n);
        maxoth = std::max(maxoth, d.noth);
    }
    const size_t per = (size_t)(kMaxOrtho + 1) * kRedBlocks;
    int rc = ws_reserve(ctx, 5, (size_t)n * per * sizeof(double));
    if (rc) return rc;
    double *partial = (double *)ctx->ws[5].ptr;
    size_t off = 0;
    const NoDesc *dd;
    if ((rc = stage(ctx, v.data(), v.size(), off, &dd))) return rc;
    X2G_LAUNCH(ctx, k_no_dot, dim3(kRedBlocks, n, (1 + maxoth + kNoGroup - 1) / kNoGroup), dim3(kRedBlock), 0, dd,
               partial);
    X2G_HIP(hipGetLastError());
    X2G_LAUNCH(ctx, k_no_apply, dim3(gx(maxn, kRedBlock), n), dim3(kRedBlock), 0, dd, partial, kRedBlocks);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_eahf_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_eahf_item *items, int n)
{
    X2G_CHECK(ctx && grids && (items || n == 0) && n >= 0, "x2g_eahf_batch: null argument");
    if (n == 0) return X2G_OK;
    size_t padtot = 0;
    std::vector<size_t> poff(n);
    for (int i = 0; i < n; ++i) {
        const x2g_eahf_item &it = items[i];
        int rc = check_grids(ctx, grids, it.grid_index);
        if (rc) return rc;
        X2G_CHECK(it.psi_d && it.cpot_d && it.xpot_d && it.ee_out_d, "x2g_eahf_batch: item %d incomplete", i);
        X2G_CHECK(it.isym == 1 || it.isym == -1, "x2g_eahf_batch: item %d isym", i);
        poff[i] = padtot;
        padtot += (grids[it.grid_index]->npad() * sizeof(double) + 255) & ~size_t(255);
    }
    int rc = ws_reserve(ctx, 4, padtot);
    if (!rc) rc = ws_reserve(ctx, 5, (size_t)n * 4 * kRedBlocks * sizeof(double));
    if (rc) return rc;
    char *pads = (char *)ctx->ws[4].ptr;
    double *partial = (double *)ctx->ws[5].ptr;
    std::vector<PadItem> pin(n);
    std::vector<EaDesc> v(n);
    for (int i = 0; i < n; ++i) {
        const x2g_eahf_item &it = items[i];
        x2g_grid *g = grids[it.grid_index];
        double *w = (double *)(pads + poff[i]);
        pin[i] = {it.psi_d, w, g->nni, g->nmu, it.isym};
        EaDesc &d = v[i];
        OneelArgs &A = d.a;
        A = OneelArgs{};
        A.w = w;
        A.psi1 = it.psi_d;
        A.psi2 = it.psi_d;
        A.wgt1 = g->suppl + 12 * g->n();
        A.e = g->suppl + 3 * g->n();
        A.f0 = g->suppl + 4 * g->n();
        A.bmu = g->bmu_orb;
        A.dnu = g->dnu;
        deriv_weights(g->coef[22], A.a1n, A.a2n);
        deriv_weights(g->coef[23], A.a1m, A.a2m);
        A.m2 = (double)(it.m * it.m);
        A.nni = g->nni;
        A.nmu = g->nmu;
        d.cpot = it.cpot_d;
        d.xpot = it.xpot_d;
        d.wgt2 = g->suppl + 13 * g->n();
        d.ee_out = it.ee_out_d;
    }
    if ((rc = launch_putin_items(ctx, pin))) return rc;
    size_t off = 0;
    const EaDesc *dd;
    if ((rc = stage(ctx, v.data(), v.size(), off, &dd))) return rc;
    X2G_LAUNCH(ctx, k_eahf, dim3(kRedBlocks, n), dim3(kRedBlock), 0, dd, partial);
    X2G_HIP(hipGetLastError());
    X2G_LAUNCH(ctx, k_eahf_finish, dim3(n), dim3(kRedBlock), 0, dd, partial, kRedBlocks);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_moments_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_moment_item *items, int n, int mpole,
                                 double *out_h)
{
    X2G_CHECK(ctx && grids && (items || n == 0) && n >= 0 && out_h, "x2g_moments_batch: null argument");
    X2G_CHECK(mpole >= 2 && mpole <= 8, "x2g_moments_batch: mpole must be 2..8");
    if (n == 0) return X2G_OK;
    std::vector<MomArgs> v(n);
    for (int i = 0; i < n; ++i) {
        const x2g_moment_item &it = items[i];
        int rc = check_grids(ctx, grids, it.grid_index);
        if (rc) return rc;
        X2G_CHECK(it.psi1_d && it.psi2_d, "x2g_moments_batch: item %d incomplete", i);
        X2G_CHECK(it.dm >= 0 && it.dm <= 4, "x2g_moments_batch: item %d |Delta m| must be 0..4", i);
        x2g_grid *g = grids[it.grid_index];
        MomArgs &A = v[i];
        A = MomArgs{};
        A.psi1 = it.psi1_d;
        A.psi2 = it.psi2_d;
        A.wgt2 = g->suppl + 13 * g->n();
        A.f4 = g->suppl + 8 * g->n();
        A.vxi = g->vxi;
        A.vxisq = g->vxisq;
        A.veta = g->veta;
        A.vetasq = g->vetasq;
        A.r2 = 0.5 * g->coef[24];
        A.nni = g->nni;
        A.nmu = g->nmu;
        A.dm = it.dm;
        A.mpole = mpole;
    }
    const int nv = 8 * n;
    int rc = ws_reserve(ctx, 5, (size_t)nv * (kRedBlocks + 1) * sizeof(double));
    if (!rc) rc = pinned_out_reserve(ctx, (size_t)nv * sizeof(double));
    if (rc) return rc;
    double *partial = (double *)ctx->ws[5].ptr;
    double *dout = partial + (size_t)nv * kRedBlocks;
    size_t off = 0;
    const MomArgs *dd;
    if ((rc = stage(ctx, v.data(), v.size(), off, &dd))) return rc;
    X2G_LAUNCH(ctx, k_moments_batch, dim3(kRedBlocks, n), dim3(kRedBlock), 0, dd, partial);
    X2G_HIP(hipGetLastError());
    X2G_LAUNCH(ctx, k_finish, dim3(nv), dim3(kRedBlock), 0, partial, kRedBlocks, nv, dout);
    X2G_HIP(hipGetLastError());
    X2G_HIP(hipMemcpyAsync(ctx->pinned_out, dout, nv * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    const double *r = (const double *)ctx->pinned_out;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 8; ++k) out_h[8 * i + k] = k < mpole ? r[8 * i + k] : 0.0;
    return X2G_OK;
}

extern "C" int x2g_pot_tail_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_tail_item *items, int n, int mpole)
{
    X2G_CHECK(ctx && grids && (items || n == 0) && n >= 0, "x2g_pot_tail_batch: null argument");
    X2G_CHECK(mpole >= 2 && mpole <= 8, "x2g_pot_tail_batch: mpole must be 2..8");
    if (n == 0) return X2G_OK;
    std::vector<TailArgs> v(n);
    int maxnni = 0;
    for (int i = 0; i < n; ++i) {
        const x2g_tail_item &it = items[i];
        int rc = check_grids(ctx, grids, it.grid_index);
        if (rc) return rc;
        X2G_CHECK(it.pot_d, "x2g_pot_tail_batch: item %d has no potential", i);
        X2G_CHECK(it.kind == 0 || it.kind == 1, "x2g_pot_tail_batch: item %d kind", i);
        X2G_CHECK(it.dm >= 0, "x2g_pot_tail_batch: item %d dm < 0", i);
        x2g_grid *g = grids[it.grid_index];
        TailArgs &A = v[i];
        A = TailArgs{};
        A.pot = it.pot_d;
        A.vxi = g->vxi;
        A.vxisq = g->vxisq;
        A.veta = g->veta;
        A.vetasq = g->vetasq;
        for (int k = 0; k < 8; ++k) A.mom[k] = k < mpole ? it.mom[k] : 0.0;
        A.r2 = 0.5 * g->coef[24];
        A.nni = g->nni;
        A.nmu = g->nmu;
        A.kind = it.kind;
        A.dm = it.dm;
        A.mpole = mpole;
        maxnni = std::max(maxnni, g->nni);
    }
    size_t off = 0;
    const TailArgs *dd;
    int rc = stage(ctx, v.data(), v.size(), off, &dd);
    if (rc) return rc;
    X2G_LAUNCH(ctx, k_pot_tail_batch, dim3(gx(4L * maxnni, 128), n), dim3(128), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

// ---------------------------------------------------------------------------
// SCF control on the device: one thread per job advances the job's control
// record by one iteration.  The rules are those of the Fortran driver's
// energy_change / convergence_control / scf_control (the reference's
// doSCF.F90:546-744 with the failure guards and the FP64 finish), in the
// same order of comparisons and assignments, so every decision is the host
// rules' decision bit for bit (only comparisons, one subtraction per orbital
// and one division; no contraction is formed).  A workgroup per job copies
// the record into LDS with all its lanes, one lane applies the rules there
// (a private copy with indexed arrays would live in scratch memory), and
// the lanes write it back.
namespace x2g {
struct CtrlDesc {
    x2g_scf_ctrl *c;
    const double *rec;
};

__device__ static bool ctrl_rules(x2g_scf_ctrl &c)
{
#pragma clang fp contract(off)
    const double eps0 = 2.220446049250313e-16;  // epsilon(1.0_dp)
    const int norb = c.norb;
    c.deltaeemax = 0.0;
    for (int i = 0; i < norb; ++i)
        if (fabs(c.deltaee[i]) >= fabs(c.deltaeemax)) c.deltaeemax = c.deltaee[i];
    c.deltanormmax = 0.0;
    for (int i = 0; i < norb; ++i)
        if (c.ifixorb[i] == 0 && fabs(c.orbnorm[i] - 1.0) >= c.deltanormmax) {
            c.deltanormmax = fabs(c.orbnorm[i] - 1.0);
            if (fabs(c.deltanormmax) < eps0) c.deltanormmax = eps0;
        }
    if (fabs(c.eeconv) < eps0) c.eeconv = c.deltaeemax;

    // failure guards: a NaN norm or energy, or divergence
    for (int i = 0; i < norb; ++i)
        if (isnan(c.ee[i]) || isnan(c.orbnorm[i]) || isnan(c.deltaee[i])) {
            c.stopcode = 7;
            return true;
        }
    if (fabs(c.deltaeemax) > c.inc_thld) {
        c.stopcode = 6;
        return true;
    }

    // saturation: no decrease for nenlast / nnolast iterations after nscf2skip
    if (c.iscf > c.nscf2skip) {
        if (fabs(c.deltaeemax) >= fabs(c.eeconvprev)) {
            ++c.noenergydec;
        } else {
            c.eeconvprev = c.deltaeemax;
            c.noenergydec = 0;
        }
        if (fabs(c.deltanormmax) >= fabs(c.dnmaxprev)) {
            ++c.nonormdec;
        } else {
            c.dnmaxprev = c.deltanormmax;
            c.nonormdec = 0;
        }
        if (c.noenergydec > c.nenlast || c.nonormdec > c.nnolast) {
            c.stopcode = 4;
            return true;
        }
    } else {
        c.eeconvprev = 100.0;
        c.dnmaxprev = c.deltanormmax;
    }
    if (c.iscf == c.maxscf) {
        c.stopcode = 3;
        return true;
    }

    // the frozen starting orbitals are released once the energies settle
    if (c.lfixorb4init && fabs(c.deltaeemax) < c.thren4init) {
        if (c.norbenergythld == c.nscfextra) {
            for (int i = 0; i < X2G_CTRL_MAXO; ++i) c.ifixorb[i] = 0;
            c.lfixorb = 0;
            c.lfixorb4init = 0;
            c.eeconv = 10.0;
            c.deltaeemax = 100.0;
        } else {
            ++c.norbenergythld;
        }
    }

    if (!c.lenergythld) {
        if (fabs(c.deltaeemax) < c.ethld) {
            c.lenergythld = 1;
            c.eeconv = c.deltaeemax;
            for (int i = 0; i < X2G_CTRL_MAXO; ++i) c.itouch[i] = 0;
            c.norbenergythld = 0;
        }
    } else {
        if (fabs(c.deltaeemax) < c.ethld) {
            if (c.norbenergythld == c.nscfextra) {
                c.stopcode = 1;
                return true;
            }
            ++c.norbenergythld;
        } else if (fabs(c.deltanormmax) < c.nthld) {
            if (c.norbnormthld == c.nscfextra) {
                c.stopcode = 2;
                return true;
            }
            ++c.norbnormthld;
        } else {
            c.lenergythld = 0;
            if (c.nresets > c.maxresets) {
                c.stopcode = 5;
                return true;
            }
            ++c.nresets;
        }
    }

    if (!c.lfixorb) {
        if (!c.lnormthld) {
            if (fabs(c.deltanormmax) < c.nthld) {
                c.lnormthld = 1;
                for (int i = 0; i < X2G_CTRL_MAXO; ++i) c.itouch[i] = 0;
                c.norbnormthld = 0;
            }
        } else {
            if (fabs(c.deltanormmax) < c.nthld) {
                if (c.norbnormthld == c.nscfextra) {
                    c.stopcode = 2;
                    return true;
                }
                ++c.norbnormthld;
            } else {
                c.lnormthld = 0;
                if (c.nresets4no > c.maxresets) {
                    c.stopcode = 5;
                    return true;
                }
                ++c.nresets4no;
            }
        }
    }

    // refresh the multipole moments when the energy change dropped enough
    if (fabs(c.eeconv / c.deltaeemax) > c.recalc_mm) {
        c.eeconv = c.deltaeemax;
        c.mm_due = 1;
        for (int i = 0; i < X2G_CTRL_MAXO; ++i) c.mm_touch[i] = c.itouch[i];
    }
    for (int i = 0; i < X2G_CTRL_MAXO; ++i) c.itouch[i] = 0;
    return false;
}

__device__ static void ctrl_step(x2g_scf_ctrl &c, const double *rec)
{
#pragma clang fp contract(off)
    c.norb = max(0, min(c.norb, X2G_CTRL_MAXO));
    // the iteration: every orbital was touched; the new norms (of the
    // orbitals not frozen) and energies, and their changes
    ++c.iscf;
    c.mm_due = 0;
    for (int i = 0; i < c.norb; ++i) {
        c.itouch[i] = 1;
        if (c.ifixorb[i] == 0) c.orbnorm[i] = rec[i];
        c.ee[i] = rec[X2G_CTRL_MAXO + i];
        c.deltaee[i] = c.ee[i] - c.eeold[i];
        if (fabs(c.deltaee[i]) < 2.220446049250313e-16) c.deltaee[i] = 2.0 * c.ethld;
        c.eeold[i] = c.ee[i];
    }
    bool stop = ctrl_rules(c);
    // the FP64 finish of a mixed-precision job
    if (c.polish > 0) {
        const bool threshold_stop = stop && c.stopcode != 3 && c.stopcode < 6;
        if (c.prec == 1) {
            if (threshold_stop || (!stop && c.iscf >= c.maxscf - c.polish)) {
                c.prec = 0;
                c.polish_end = min(c.iscf + c.polish, c.maxscf);
                c.lenergythld = 0;
                c.lnormthld = 0;
                c.norbenergythld = 0;
                c.norbnormthld = 0;
                c.nresets = 0;
                c.nresets4no = 0;
                c.noenergydec = 0;
                c.nonormdec = 0;
                c.eeconvprev = 1.0e35;
                c.dnmaxprev = 1.0e35;
                c.stopcode = 0;
                stop = false;
            }
        } else if (threshold_stop && c.iscf < c.polish_end) {
            c.stopcode = 0;
            stop = false;
        }
    }
    c.stop = stop ? 1 : 0;
}

constexpr int kCtrlWords = (int)(sizeof(x2g_scf_ctrl) / sizeof(int));
static_assert(sizeof(x2g_scf_ctrl) % sizeof(double) == 0, "x2g_scf_ctrl: whole doubles");

__global__ void __launch_bounds__(64) k_scf_control(const CtrlDesc *__restrict__ cd)
{
    __shared__ x2g_scf_ctrl sc;
    __shared__ double srec[2 * X2G_CTRL_MAXO];
    const CtrlDesc d = cd[blockIdx.x];
    int *w = reinterpret_cast<int *>(&sc);
    const int *src = reinterpret_cast<const int *>(d.c);
    for (int i = threadIdx.x; i < kCtrlWords; i += blockDim.x) w[i] = src[i];
    if (threadIdx.x < 2 * X2G_CTRL_MAXO) srec[threadIdx.x] = d.rec[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) ctrl_step(sc, srec);
    __syncthreads();
    int *dst = reinterpret_cast<int *>(d.c);
    for (int i = threadIdx.x; i < kCtrlWords; i += blockDim.x) dst[i] = w[i];
}
}  // namespace x2g

extern "C" int x2g_scf_control(x2g_ctx *ctx, x2g_scf_ctrl *const *ctrl_d, const double *const *rec_d, int n)
{
    X2G_CHECK(ctx && n >= 0 && (n == 0 || (ctrl_d && rec_d)), "x2g_scf_control: null argument");
    if (n == 0) return X2G_OK;
    std::vector<x2g::CtrlDesc> v(n);
    for (int k = 0; k < n; ++k) {
        X2G_CHECK(ctrl_d[k] && rec_d[k], "x2g_scf_control: job %d has a null record", k);
        v[k] = {ctrl_d[k], rec_d[k]};
    }
    size_t off = 0;
    const x2g::CtrlDesc *dd;
    int rc = stage(ctx, v.data(), v.size(), off, &dd);
    if (rc) return rc;
    X2G_LAUNCH(ctx, x2g::k_scf_control, dim3(n), dim3(64), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

extern "C" int x2g_scf_ctrl_bytes(void) { return (int)sizeof(x2g_scf_ctrl); }

extern "C" int x2g_gather(x2g_ctx *ctx, const double *const *src_d, const int *counts, int n, double *out_h)
{
    X2G_CHECK(ctx && (n == 0 || (src_d && counts && out_h)) && n >= 0, "x2g_gather: null argument");
    if (n == 0) return X2G_OK;
    std::vector<GatherDesc> v(n);
    int tot = 0;
    for (int i = 0; i < n; ++i) {
        X2G_CHECK(src_d[i] && counts[i] >= 0, "x2g_gather: item %d", i);
        v[i] = {src_d[i], counts[i], tot};
        tot += counts[i];
    }
    int rc = ws_reserve(ctx, 1, (size_t)std::max(tot, 1) * sizeof(double));
    if (!rc) rc = pinned_out_reserve(ctx, (size_t)std::max(tot, 1) * sizeof(double));
    if (rc) return rc;
    double *dbuf = (double *)ctx->ws[1].ptr;
    size_t off = 0;
    const GatherDesc *dd;
    if ((rc = stage(ctx, v.data(), v.size(), off, &dd))) return rc;
    X2G_LAUNCH(ctx, k_gather, dim3(n), dim3(64), 0, dd, dbuf);
    X2G_HIP(hipGetLastError());
    X2G_HIP(hipMemcpyAsync(ctx->pinned_out, dbuf, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(out_h, ctx->pinned_out, (size_t)tot * sizeof(double));
    return X2G_OK;
}
