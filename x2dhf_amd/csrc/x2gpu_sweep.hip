// MCSOR relaxation of the 8th-order 17-point cross stencil on the (nu,mu)
// prolate-spheroidal grid -- MI355X (gfx950) kernels.
//
// Method (unchanged from the reference, sormcsor.F90:143-386): each sweep
// updates the relaxed interior nu=2..nni-1, mu=2..nmu-4 in five colour passes,
//   f(p) = omega*(rhs - L f)/lhs + (1-omega)*f(p),
// with L the 9-point d2/dmu2 + b(mu) d/dmu + d2/dnu2 + d(nu) d/dnu stencil
// (lhs carries the diagonal), and then extrapolates the mu=1, nu=1 and
// nu=nni boundary lines (sormcsor.F90:333-383).  The colour of (nu,mu) is
// the reference's col-wise pass index ((mu-2)*(nni-2) + (nu-2)) mod 5
// (initCBAllocArrays.F90:2467-2497 with the pass split of sormcsor.F90:170-297);
// for nni = 2 (mod 5), where that pass split is not a proper colouring, the
// geometric colouring (nu - mu) mod 5 is used instead.
//
// Two implementations:
//  * k_phase / k_boundary: one launch per colour pass, in place.  Simple,
//    used as the cross-check path and for grids too wide for the LDS ring.
//  * k_strip: one launch per sweep.  A workgroup owns a strip of full nu
//    rows [a, a+L) of one problem and pipelines the five colour passes
//    through an LDS ring of 4-row blocks (pass p runs SKEW blocks behind pass
//    p-1), recomputing a 16-row halo on each side; it reads the previous
//    sweep from `src` and writes the final rows (plus the nu=1/nu=nni boundary
//    values, and the mu=1 row for the first strip) to `dst`.  Every problem
//    of a batch, on any grid, is handled by the same launch.
//  * k_sweep: the same strip pipeline with the lhs/rhs pairs pre-split by
//    colour, compile-time ring strides and one 16-byte chunk per block copy;
//    each workgroup walks a balanced list of strips (row segments) across
//    the batch's problems.
#include "x2gpu_internal.h"
#include <algorithm>
#include <string>
#include <atomic>
#include <cstring>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <type_traits>

namespace x2g {

__host__ __device__ inline int colour_mult(int nni)
{
    int a = (nni - 2) % 5;
    return a == 0 ? 4 : a;
}

__device__ __forceinline__ int mod5(int x)
{
    int r = x % 5;
    return r < 0 ? r + 5 : r;
}

template <typename T>
struct SweepDesc {
    const T *src;
    T *dst;
    const T *lhs;
    const T *rhs;
    const T *bmu;
    const T *dnu;
    T c[16];  // dmu1[4], dmu2[4], dni1[4], dni2[4]
    T ex[5];
    T omega, omega1;
    const T *lr;          // k_sweep: (omega/lhs, rhs) pairs in the residue-split order
    unsigned pad_bytes;   // bytes of src (and of dst)
    unsigned lr_bytes;    // bytes of lr
    int P;                // row stride of src / dst (nni+8 except for k_sweep)
    int nni, nmu, isym, cA;
    int wg_begin, L;      // k_strip: first workgroup and strip height
};

// ---------------------------------------------------------------------------
// Point update, term order of sormcsor.F90:173-194.  `f` addresses the padded
// array, s1 is the mu stride (nni+8).
template <typename T>
__device__ __forceinline__ T stencil_cc(const T *f, long p, long s1, const T *c, T b, T d)
{
    const long s2 = 2 * s1, s3 = 3 * s1, s4 = 4 * s1;
    T ddmi2 = c[4] * (f[p - s4] + f[p + s4]) + c[5] * (f[p - s3] + f[p + s3]) +
              c[6] * (f[p - s2] + f[p + s2]) + c[7] * (f[p - s1] + f[p + s1]);
    T ddmi1 = c[0] * (f[p - s4] - f[p + s4]) + c[1] * (f[p - s3] - f[p + s3]) +
              c[2] * (f[p - s2] - f[p + s2]) + c[3] * (f[p - s1] - f[p + s1]);
    T ddni2 = c[12] * (f[p - 4] + f[p + 4]) + c[13] * (f[p - 3] + f[p + 3]) +
              c[14] * (f[p - 2] + f[p + 2]) + c[15] * (f[p - 1] + f[p + 1]);
    T ddni1 = c[8] * (f[p - 4] - f[p + 4]) + c[9] * (f[p - 3] - f[p + 3]) +
              c[10] * (f[p - 2] - f[p + 2]) + c[11] * (f[p - 1] - f[p + 1]);
    return ddmi2 + b * ddmi1 + ddni2 + d * ddni1;
}

// ---------------------------------------------------------------------------
// Per-pass in-place kernel: blockIdx.y = problem.
template <typename T>
__global__ void k_phase(const SweepDesc<T> *__restrict__ descs, int pass)
{
    const SweepDesc<T> &d = descs[blockIdx.y];
    const int nni = d.nni, nrow = nni - 2;
    const long s1 = nni + 8;
    const long npts = (long)nrow * (d.nmu - 5);
    T *f = d.dst;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < npts;
         k += (long)gridDim.x * blockDim.x) {
        const int mu = 2 + (int)(k / nrow), nu = 2 + (int)(k % nrow);
        if (mod5((nu - 2) + d.cA * (mu - 2)) != pass) continue;
        const long p = (long)(mu + 3) * s1 + (nu + 3);
        const long u = (long)(mu - 1) * nni + (nu - 1);
        const T cc = stencil_cc(f, p, s1, d.c, d.bmu[mu - 1], d.dnu[nu - 1]);
        f[p] = d.omega * (d.rhs[u] - cc) / d.lhs[u] + d.omega1 * f[p];
    }
}

// Boundary lines after a sweep (sormcsor.F90:94-129 with the index sets of
// initCBAllocArrays.F90:2823-2842).  The corner values at mu=1 are formed
// from the freshly extrapolated mu=1 row exactly as the reference's sequential
// order does, but recomputed per thread so no cross-thread ordering is needed.
template <typename T>
__device__ __forceinline__ T row1_value(const T *f, int nu, long s1, const T *ex)
{
    const long p = (long)(1 + 3) * s1 + (nu + 3);
    return ex[0] * f[p + s1] + ex[1] * f[p + 2 * s1] + ex[2] * f[p + 3 * s1] +
           ex[3] * f[p + 4 * s1] + ex[4] * f[p + 5 * s1];
}

template <typename T>
__global__ void k_boundary(const SweepDesc<T> *__restrict__ descs)
{
    const SweepDesc<T> &d = descs[blockIdx.y];
    const int nni = d.nni, nmu = d.nmu;
    const long s1 = nni + 8;
    T *f = d.dst;
    const int nrow1 = nni - 2;       // mu=1 row, nu=2..nni-1
    const int ncol = nmu - 4;        // nu=1 / nu=nni columns, mu=1..nmu-4
    const int total = nrow1 + 2 * ncol;
    const bool even = d.isym == 1;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        if (t < nrow1) {
            const int nu = 2 + t;
            f[(long)4 * s1 + nu + 3] = even ? row1_value(f, nu, s1, d.ex) : T(0);
        } else {
            const int q = t - nrow1;
            const int side = q / ncol;          // 0: nu=1, 1: nu=nni
            const int mu = 1 + q % ncol;
            const int nu = side == 0 ? 1 : nni;
            const int dir = side == 0 ? 1 : -1;
            T v = T(0);
            if (even) {
                if (mu == 1) {
                    v = d.ex[0] * row1_value(f, nu + dir, s1, d.ex) +
                        d.ex[1] * row1_value(f, nu + 2 * dir, s1, d.ex) +
                        d.ex[2] * row1_value(f, nu + 3 * dir, s1, d.ex) +
                        d.ex[3] * row1_value(f, nu + 4 * dir, s1, d.ex) +
                        d.ex[4] * row1_value(f, nu + 5 * dir, s1, d.ex);
                } else {
                    const long p = (long)(mu + 3) * s1 + (nu + 3);
                    v = d.ex[0] * f[p + dir] + d.ex[1] * f[p + 2 * dir] + d.ex[2] * f[p + 3 * dir] +
                        d.ex[3] * f[p + 4 * dir] + d.ex[4] * f[p + 5 * dir];
                }
            }
            f[(long)(mu + 3) * s1 + nu + 3] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Pipelined strip kernel.
//
// Rows are grouped in blocks of 4: block b of a strip starting at row a
// holds rows r0+4b .. r0+4b+3 with r0 = a-16.  Pass p updates blocks
// b in [p, K-1-p] (K = L/4 + 8), i.e. rows [a-16+4p, a+L+16-4p), and at step
// s it processes block s - SKEW*p.  Pass p of a block reads rows within one
// block of it; with SKEW = 2 the blocks touched by different passes in one
// step are two or more blocks apart, so a step needs a single barrier; with
// SKEW = 1 the five passes of a step are separated by barriers.  Colours of
// lower passes are always final and colours of higher passes still hold the
// previous sweep when a block is processed, which is exactly the sequential
// five-pass order restricted to the strip; the 16-row halo makes the rows
// [a, a+L) independent of what lies beyond the strip's 20-row read window.
//
// LDS: a ring of RB = 4*SKEW+4 blocks (4 rows of nni+8 values each); during
// step s blocks s-4*SKEW-1 .. s+1 are live and block s+2 is prefetched
// through registers while the step computes.
constexpr int kMaxLoads = 8;  // prefetch registers per thread (4*(nni+8) <= 8*blockDim)
// k_sweep copies a block with NS = 5 (one colour point per lane and row) or
// 8 (two) slots per thread: 4*(nni+8) <= NS*blockDim

template <typename T, int SKEW>
__global__ void __launch_bounds__(512) k_strip(const SweepDesc<T> *__restrict__ descs, int nprob)
{
    extern __shared__ unsigned char smem_raw[];
    T *ring = reinterpret_cast<T *>(smem_raw);
    constexpr int RB = 4 * SKEW + 4;

    // locate the problem owning this workgroup (descs sorted by wg_begin)
    const int wg = blockIdx.x;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (descs[mid].wg_begin <= wg) lo = mid; else hi = mid - 1;
    }
    const SweepDesc<T> &d = descs[lo];
    const int nni = d.nni, nmu = d.nmu, N8 = nni + 8, L = d.L;
    const int chunk = wg - d.wg_begin;
    const int a = 2 + chunk * L;
    const int aend = min(a + L, nmu - 3);
    const int r0 = a - 16;
    const int K = L / 4 + 8;
    const int S = K + 4 * SKEW - 3;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int M = (nni - 2 + 4) / 5;      // max colour points per row
    const bool even = d.isym == 1;
    const T *__restrict__ src = d.src;
    T *__restrict__ dst = d.dst;

    auto lds_row = [&](int mu) -> T * {
        const int q = mu - r0 + 4;          // >= 0 for every row touched
        const int b = (q >> 2) - 1;
        const int slot = (b + RB) % RB;
        return ring + (long)(slot * 4 + (q & 3)) * N8;
    };
    auto load_block_direct = [&](int b) {
        const int slot = (b + RB) % RB;
        for (int e = tid; e < 4 * N8; e += nth) {
            const int rr = e / N8, col = e - rr * N8;
            const int mu = r0 + 4 * b + rr;
            T v = T(0);
            if (mu >= -3 && mu <= nmu + 4) v = src[(long)(mu + 3) * N8 + col];
            ring[(long)(slot * 4 + rr) * N8 + col] = v;
        }
    };

    load_block_direct(-1);
    load_block_direct(0);
    load_block_direct(1);
    __syncthreads();

    for (int s = 0; s < S; ++s) {
        // 1. prefetch block s+2 into registers
        const int bl = s + 2;
        T pre[kMaxLoads];
        const bool do_load = bl <= K;
        if (do_load) {
#pragma unroll
            for (int j = 0; j < kMaxLoads; ++j) {
                const int e = tid + j * nth;
                T v = T(0);
                if (e < 4 * N8) {
                    const int rr = e / N8, col = e - rr * N8;
                    const int mu = r0 + 4 * bl + rr;
                    if (mu >= -3 && mu <= nmu + 4) v = src[(long)(mu + 3) * N8 + col];
                }
                pre[j] = v;
            }
        }

        // 2. write out the block finalised by the last pass in step s-1
        const int bw = s - 4 * SKEW - 1;
        if (bw >= 4 && bw <= K - 5) {
            for (int e = tid; e < 4 * nni; e += nth) {
                const int rr = e / nni, nu = 1 + e - rr * nni;
                const int mu = r0 + 4 * bw + rr;
                if (mu < a || mu >= aend) continue;
                const T *row = lds_row(mu);
                T v;
                if (nu == 1 || nu == nni) {
                    const int dir = nu == 1 ? 1 : -1;
                    const int c0 = nu + 3;
                    v = even ? d.ex[0] * row[c0 + dir] + d.ex[1] * row[c0 + 2 * dir] +
                                   d.ex[2] * row[c0 + 3 * dir] + d.ex[3] * row[c0 + 4 * dir] +
                                   d.ex[4] * row[c0 + 5 * dir]
                             : T(0);
                } else {
                    v = row[nu + 3];
                }
                dst[(long)(mu + 3) * N8 + nu + 3] = v;
            }
            // the mu=1 row once rows 2..6 are final (first strip only)
            if (a == 2 && bw == 5) {
                const T *rw[5] = {lds_row(2), lds_row(3), lds_row(4), lds_row(5), lds_row(6)};
                auto r1 = [&](int nu) -> T {
                    const int c0 = nu + 3;
                    return d.ex[0] * rw[0][c0] + d.ex[1] * rw[1][c0] + d.ex[2] * rw[2][c0] +
                           d.ex[3] * rw[3][c0] + d.ex[4] * rw[4][c0];
                };
                for (int nu = 1 + tid; nu <= nni; nu += nth) {
                    T v = T(0);
                    if (even) {
                        if (nu == 1 || nu == nni) {
                            const int dir = nu == 1 ? 1 : -1;
                            v = d.ex[0] * r1(nu + dir) + d.ex[1] * r1(nu + 2 * dir) +
                                d.ex[2] * r1(nu + 3 * dir) + d.ex[3] * r1(nu + 4 * dir) +
                                d.ex[4] * r1(nu + 5 * dir);
                        } else {
                            v = r1(nu);
                        }
                    }
                    dst[(long)4 * N8 + nu + 3] = v;
                }
            }
        }

        // 3. the five colour passes of this step
        for (int pp = 0; pp < (SKEW == 1 ? 5 : 1); ++pp) {
            const int p_lo = SKEW == 1 ? pp : 0, p_hi = SKEW == 1 ? pp + 1 : 5;
            const int nitems = (p_hi - p_lo) * 4 * M;
            for (int it = tid; it < nitems; it += nth) {
                const int p = p_lo + it / (4 * M);
                const int rr = (it / M) & 3;
                const int t = it % M;
                const int b = s - SKEW * p;
                if (b < p || b > K - 1 - p) continue;
                const int mu = r0 + 4 * b + rr;
                if (mu < 2 || mu > nmu - 4) continue;
                const int nu = 2 + mod5(p - d.cA * (mu - 2)) + 5 * t;
                if (nu > nni - 1) continue;
                const int c0 = nu + 3;
                T *rc = lds_row(mu);
                const T *m1 = lds_row(mu - 1), *m2 = lds_row(mu - 2), *m3 = lds_row(mu - 3),
                        *m4 = lds_row(mu - 4);
                const T *p1 = lds_row(mu + 1), *p2 = lds_row(mu + 2), *p3 = lds_row(mu + 3),
                        *p4 = lds_row(mu + 4);
                const T ddmi2 = d.c[4] * (m4[c0] + p4[c0]) + d.c[5] * (m3[c0] + p3[c0]) +
                                d.c[6] * (m2[c0] + p2[c0]) + d.c[7] * (m1[c0] + p1[c0]);
                const T ddmi1 = d.c[0] * (m4[c0] - p4[c0]) + d.c[1] * (m3[c0] - p3[c0]) +
                                d.c[2] * (m2[c0] - p2[c0]) + d.c[3] * (m1[c0] - p1[c0]);
                const T ddni2 = d.c[12] * (rc[c0 - 4] + rc[c0 + 4]) + d.c[13] * (rc[c0 - 3] + rc[c0 + 3]) +
                                d.c[14] * (rc[c0 - 2] + rc[c0 + 2]) + d.c[15] * (rc[c0 - 1] + rc[c0 + 1]);
                const T ddni1 = d.c[8] * (rc[c0 - 4] - rc[c0 + 4]) + d.c[9] * (rc[c0 - 3] - rc[c0 + 3]) +
                                d.c[10] * (rc[c0 - 2] - rc[c0 + 2]) + d.c[11] * (rc[c0 - 1] - rc[c0 + 1]);
                const T cc = ddmi2 + d.bmu[mu - 1] * ddmi1 + ddni2 + d.dnu[nu - 1] * ddni1;
                const long u = (long)(mu - 1) * nni + (nu - 1);
                rc[c0] = d.omega * (d.rhs[u] - cc) / d.lhs[u] + d.omega1 * rc[c0];
            }
            if (SKEW == 1 && pp < 4) __syncthreads();
        }

        // 4. commit the prefetched block.  Its slot held block s-4*SKEW-2,
        // which only the mu=1 row of step 2 (first strip, bw == 5) still reads:
        // wait for those reads before overwriting (a, bw are uniform).
        if (a == 2 && bw == 5) __syncthreads();
        if (do_load) {
            const int slot = (bl + RB) % RB;
#pragma unroll
            for (int j = 0; j < kMaxLoads; ++j) {
                const int e = tid + j * nth;
                if (e < 4 * N8) {
                    const int rr = e / N8, col = e - rr * N8;
                    ring[(long)(slot * 4 + rr) * N8 + col] = pre[j];
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Colour-ordered layout of lhs / rhs for k_sweep.  Relaxed point (nu, mu),
// nu = 2..nni-1, mu = 2..nmu-4, of colour c = (nu-2 + cA*(mu-2)) mod 5 is
// stored at
//   (mu-2)*5*M + c*M + (nu-2) div 5,   M = ceil((nni-2)/5),
// so the points of one colour pass of one row are contiguous and the t-th
// point of pass c in any row sits at position t of that row's c-block (a row
// holds M or M-1 points of each colour; a short block leaves its last slot
// unused).
__host__ __device__ inline int split_m(int nni) { return (nni - 2 + 4) / 5; }
__host__ __device__ inline long split_size(int nni, int nmu) { return (long)(nmu - 5) * 5 * split_m(nni); }
__device__ __forceinline__ long split_index(int nu, int mu, int M, int cA)
{
    const int q = nu - 2;
    return (long)(mu - 2) * 5 * M + (long)((q + cA * (mu - 2)) % 5) * M + q / 5;
}

template <typename TO>
struct PackDesc {
    const double *src;  // unpadded nni*nmu array
    TO *dst;            // residue-split array of pairs: component `comp` of pair i
    double omega;       // > 0: store omega/src (the scaled inverse diagonal), else src
    int nni, nmu, comp;
};

template <typename TO>
__global__ void k_pack(const PackDesc<TO> *__restrict__ pd)
{
    const PackDesc<TO> &d = pd[blockIdx.y];
    const int nni = d.nni, nrow = nni - 2, M = split_m(nni), cA = colour_mult(nni);
    const long npts = (long)nrow * (d.nmu - 5);
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < npts; k += (long)gridDim.x * blockDim.x) {
        const int mu = 2 + (int)(k / nrow), nu = 2 + (int)(k % nrow);
        const double v = d.src[(long)(mu - 1) * nni + (nu - 1)];
        d.dst[2 * split_index(nu, mu, M, cA) + d.comp] = (TO)(d.omega > 0.0 ? d.omega / v : v);
    }
}

// rhs = f[0] * f[1] * f[2] (null f[1] / f[2] count as 1) for the items that
// give their right-hand side as a product and relax in FP64.
struct ProdDesc {
    double *out;
    const double *f[3];
    long n;
};
__global__ void k_rhs_product(const ProdDesc *__restrict__ pd)
{
    const ProdDesc &d = pd[blockIdx.y];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (long)gridDim.x * blockDim.x) {
        double v = d.f[0][e];
        if (d.f[1]) v *= d.f[1][e];
        if (d.f[2]) v *= d.f[2][e];
        d.out[e] = v;
    }
}

// ---------------------------------------------------------------------------
// k_sweep: the pipelined strip sweep (the block schedule of k_strip, with
// blocks of BR = 4 or 2 rows) with a fixed lane -> work map.  Wave p < 5
// serves colour pass p: its lane l the colour points t = l + 64 j (j < NT) of
// each of the BR rows of the pass's current block.  With the pass uniform across the wave, all
// that depends on it (block, ring slot, rows, colour offset, lhs / rhs row
// base) is scalar work.  The points t >= 64 NT of each pass (MS of them per
// row at most) go to spill waves, one point per lane, pass (tid-320) / MS.
// The ring rows have a compile-time
// stride W >= nni+16, so a point reads its 17 stencil values with immediate
// offsets from 2 (4/BR) + 1 block bases, and every point of a step is processed
// without branches: points outside the pass's block range or the grid are
// computed on in-bounds garbage and written to an LDS sink.
//
// Global memory goes through buffer resources (raw buffers: a load at an
// offset outside the array returns 0 and a store there is dropped).  Every
// lane issues every load and store with a 32-bit offset, inactive lanes with
// an out-of-range one, so each wave issues the same memory operations in
// every step and the compiler's waits count exactly the loads they need.
// src / dst have a row stride P = nni+8 rounded up to 16 bytes: a block of
// BR rows is a contiguous run of 16-byte chunks, copied into the ring with one
// 16-byte load and one LDS store per chunk (NSV chunks per thread) and
// written back the same way, pad columns included (src and dst hold the same
// pads, which the sweep never changes).  lhs / rhs come as one array of
// (omega/lhs, rhs) pairs in the residue-split order (contiguous per pass and
// row, the t-th point of a pass at position t), one pair load per point.
// A point update is
//   f = (rhs - L f) * (omega/lhs) + (1-omega) f,
// the d/dmu part of L f with per-row weights  c2_k +- b(mu) c1_k  (k = 1..4);
// FP32 sums L f in one fused multiply-add chain (28 operations per point).
// The rows of a block are taken two at a time: the 2 x 17 stencil values are
// read from LDS before the arithmetic of either row.  The next block and the
// next step's (omega/lhs, rhs) are loaded one step ahead.  The mu = 1 row of
// the first strip is formed after the last step from the written rows 2..6,
// so the steps issue a fixed number of memory operations.
//
// LDS (elements of T): d(nu) at index nu-1 [W] (so a point's column base also
// addresses its d(nu)), the row weights of the ring's rows [8 per row, written
// with each block], the coefficients [32], the ring [its rows x W], the sink
// [one per thread].  Independent of the strip height.  Two-row blocks halve
// the per-lane registers of a step (the (omega/lhs, rhs) pairs and the
// prefetched chunks), so more workgroups share a CU, at the price of a
// longer pipeline fill per strip and two more ring rows per pass skew.
//
// Pipeline shape of k_sweep for blocks of BR rows (4 or 2): passes SKEW
// blocks apart (at least one block more than the stencil's reach of 4 rows),
// and the ring holds the live window of 4 SKEW + 2 (4/BR) + 1 blocks plus the
// one being prefetched: 12 blocks of 4 rows (48 rows) or 18 of 2 (36 rows).
__host__ __device__ constexpr int sweep_skew(int BR) { return BR == 4 ? 2 : 3; }
__host__ __device__ constexpr int sweep_ring_blocks(int BR) { return 4 * sweep_skew(BR) + 2 * (4 / BR) + 2; }
// LDS of k_sweep (elements of T): d(nu) [W], the row weights [8 per ring
// row], the coefficients [32], the ring [ring rows x W], the sink [nth]
__host__ __device__ inline size_t sweep_lds_elems(int W, int nth, int BR)
{
    const int rows = sweep_ring_blocks(BR) * BR;
    return (size_t)W + (size_t)rows * 8 + 32 + (size_t)rows * W + (size_t)nth;
}

// elements of T per 16-byte chunk, and chunks per thread in a block copy (a
// block is 4P elements with P <= nni+11, and a launch has at least
// max(64, nni-2) threads)
template <typename T>
constexpr int sweep_vec() { return 16 / (int)sizeof(T); }
template <typename T>
constexpr int sweep_nsv(int BR) { return sizeof(T) == 4 ? (BR == 4 ? 2 : 1) : (BR == 4 ? 3 : 2); }
// row stride of k_sweep's padded arrays: nni+8 rounded up to 16 bytes
__host__ __device__ inline int sweep_row_stride(int nni, int vec) { return (nni + 8 + vec - 1) / vec * vec; }

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// Byte offsets outside every array (arrays are < 1 GiB): kOOB for a row or
// lane without data, kOOBLane added to a row offset for an absent point (the
// sum of both stays below 2^32).
constexpr unsigned kOOB = 0x80000000u;
constexpr unsigned kOOBLane = 0x40000000u;
constexpr int kRsrcWord3 = 0x00020000;  // buffer resource word 3 for raw buffers on gfx9-family GPUs

// Byte offset `off` where ok, else one outside every array (an offset with
// the top bit set; `off` is formed either way, which keeps the address
// arithmetic free of branches).
__device__ __forceinline__ unsigned oob_unless(bool ok, int off) { return (unsigned)off | (ok ? 0u : kOOB); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kRsrcWord3);
}
__device__ __forceinline__ float buf_load1(__amdgpu_buffer_rsrc_t r, unsigned off, float)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ double buf_load1(__amdgpu_buffer_rsrc_t r, unsigned off, double)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store1(__amdgpu_buffer_rsrc_t r, unsigned off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_store1(__amdgpu_buffer_rsrc_t r, unsigned off, double v)
{
    typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, off, 0, 0);
}
template <typename T>
struct Pair {
    T a, b;
};
__device__ __forceinline__ Pair<float> buf_load2(__amdgpu_buffer_rsrc_t r, unsigned off, float)
{
    return __builtin_bit_cast(Pair<float>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ Pair<double> buf_load2(__amdgpu_buffer_rsrc_t r, unsigned off, double)
{
    return __builtin_bit_cast(Pair<double>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
template <typename T>
struct Chunk {
    T v[16 / sizeof(T)];
};

// A k_sweep launch runs a list of row segments per workgroup: wg_seg[w] ..
// wg_seg[w+1]-1 index segs, each strip [a, a+L) of problem prob.
struct SweepSeg {
    int prob, a, L;
};
template <typename T, int NT, int W, int BR>
__device__ __forceinline__ void sweep_segment(const SweepDesc<T> &d, const int a, const int L, const int MS,
                                              const bool zsrc, unsigned char *ks_smem);

// Waves per SIMD the register allocation leaves room for, from the
// workgroups the LDS lets share a CU: FP32 rings of stride <= 384 fit two
// 6-wave workgroups (4-row blocks), with 2-row blocks up to stride 256 three
// and more, stride 384 two and stride 528 two of 8 waves.
template <typename T, int W, int BR>
constexpr int sweep_waves_per_eu()
{
    return sizeof(T) != 4 ? 1 : (BR == 4 ? (W <= 384 ? 4 : 1) : (W <= 256 ? 5 : (W <= 528 ? 4 : 1)));
}

template <typename T, int NT, int W, int BR>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sweep_waves_per_eu<T, W, BR>())))
k_sweep(const SweepDesc<T> *__restrict__ descs, const SweepSeg *__restrict__ segs, const int *__restrict__ wg_seg,
        int MS, int zero_src)
{
    extern __shared__ __align__(16) unsigned char ks_smem[];
    const int s0 = wg_seg[blockIdx.x], s1 = wg_seg[blockIdx.x + 1];
    for (int si = s0; si < s1; ++si) {
        const SweepSeg sg = segs[si];
        if (si > s0) __syncthreads();  // the previous segment is done with the LDS
        sweep_segment<T, NT, W, BR>(descs[sg.prob], sg.a, sg.L, MS, zero_src != 0, ks_smem);
    }
}

// One strip of rows [a, a+L) of problem d (L a multiple of BR, or the rest of
// the problem): the pipelined sweep described above, with blocks of BR rows
// (4 or 2).  zsrc: src is known to be zero everywhere (the first sweep of a
// correction), so it is not read.
template <typename T, int NT, int W, int BR>
__device__ __forceinline__ void sweep_segment(const SweepDesc<T> &d, const int a, const int L, const int MS,
                                              const bool zsrc, unsigned char *ks_smem)
{
    // R: stencil reach in blocks; SKEW: blocks between consecutive passes
    // (>= 1 + R, so one barrier per step suffices); RB: ring blocks (the
    // live window s-4 SKEW-R .. s+R plus the block being prefetched)
    constexpr int R = 4 / BR, SKEW = sweep_skew(BR), RB = sweep_ring_blocks(BR), RROWS = RB * BR;
    constexpr int VEC = sweep_vec<T>(), NSV = sweep_nsv<T>(BR);
    T *const dnu_s = reinterpret_cast<T *>(ks_smem);

    const int nni = d.nni, nmu = d.nmu, P = d.P;
    const int aend = min(a + L, nmu - 3);
    const int r0 = a - 16;
    const int K = (L + 32) / BR;          // blocks of the strip and its 16-row halos
    const int S = K + 4 * SKEW - 3 * R;   // steps: the last block's write-out at K-1-4R + 4 SKEW + R
    const int tid = threadIdx.x, nth = blockDim.x;
    const int M = split_m(nni);
    const int cA = d.cA;
    const bool even = d.isym == 1;
    const __amdgpu_buffer_rsrc_t rs_src = buf_rsrc(d.src, d.pad_bytes);
    const __amdgpu_buffer_rsrc_t rs_dst = buf_rsrc(d.dst, d.pad_bytes);
    const __amdgpu_buffer_rsrc_t rs_lr = buf_rsrc(d.lr, d.lr_bytes);
    const __amdgpu_buffer_rsrc_t rs_bmu = buf_rsrc(d.bmu, (unsigned)(nmu * sizeof(T)));
    T *const ab_r = dnu_s + W;                 // row weights, 8 per ring row
    T *const coef_s = ab_r + RROWS * 8;        // c[8..15] (nu weights), 1-omega
    T *const ring = coef_s + 32;
    T *const sink = ring + RROWS * W;

    auto slot_of = [&](int b) { return (b + RB) % RB; };  // b >= -RB

    // Chunk q = tid + j*nth of a block (elements VEC*q .. VEC*q+VEC-1 of its
    // BR x P values): its LDS offset rr*W + col within a ring block (-1: no
    // such chunk), and in cinfo its row rr (bits 0..2; BR: none), the element
    // of the chunk holding nu = 1 (bits 3..5) and nu = nni (bits 6..8) of that
    // row (7: none).  Fixed for the whole strip.
    int cl[NSV], cinfo[NSV];
#pragma unroll
    for (int j = 0; j < NSV; ++j) {
        const int e0 = (tid + j * nth) * VEC;
        const int rr = e0 / P, col = e0 - rr * P;
        const bool v = e0 < BR * P;
        const int b1 = 4 - col, b2 = nni + 3 - col;
        cl[j] = v ? rr * W + col : -1;
        cinfo[j] = (v ? rr : BR) | ((v && b1 >= 0 && b1 < VEC ? b1 : 7) << 3) |
                   ((v && b2 >= 0 && b2 < VEC ? b2 : 7) << 6);
    }

    // Row weights of row mu: distance k = 4,3,2,1 -> c2 = c[4..7], c1 = c[0..3];
    // ab[0..3] = c2 + b(mu) c1 (rows mu-k), ab[4..7] = c2 - b(mu) c1 (rows mu+k).
    // Thread tid < 8 BR forms entry tid & 7 of row tid >> 3 of a block; its
    // two coefficients are loaded once.
    const T rw_c2 = d.c[4 + (tid & 3)], rw_c1 = d.c[tid & 3];
    auto row_weight = [&](T b) -> T { return (tid & 7) < 4 ? rw_c2 + b * rw_c1 : rw_c2 - b * rw_c1; };

    // the chunks of block bl, plus b(mu) of row (tid >> 3) % BR of the block
    // for the row weights (rows outside the grid read b = 0)
    auto load_block = [&](int bl, u32x4 (&pr)[NSV], T &bv) {
        const int mub = r0 + BR * bl;
        {
            const int mu = mub + ((tid >> 3) & (BR - 1));
            const bool ok = tid < 8 * BR && mu >= 1 && mu <= nmu;
            bv = buf_load1(rs_bmu, oob_unless(ok, (mu - 1) * (int)sizeof(T)), T(0));
        }
        const int gb = (mub + 3) * P * (int)sizeof(T);  // negative above the padded array
#pragma unroll
        for (int j = 0; j < NSV; ++j) {
            const int mu = mub + (cinfo[j] & 7);
            const bool ok = !zsrc && cl[j] >= 0 && mu >= -3 && mu <= nmu + 4;
            pr[j] = __builtin_amdgcn_raw_buffer_load_b128(rs_src, oob_unless(ok, gb + (tid + j * nth) * 16), 0, 0);
        }
    };
    auto commit_block = [&](int bl, const u32x4 (&pr)[NSV], T bv) {
        if (bl > K - 1 + R) return;
        const int sl = slot_of(bl);
        if (tid < 8 * BR) ab_r[sl * 8 * BR + tid] = row_weight(bv);
        T *bp = ring + sl * BR * W;
#pragma unroll
        for (int j = 0; j < NSV; ++j)
            if (cl[j] >= 0) *reinterpret_cast<u32x4 *>(bp + cl[j]) = pr[j];
    };
    // rows [a, aend) of block bw (final) to dst, with the nu = 1 and nu = nni
    // values extrapolated (sormcsor.F90:352-383; zero for odd symmetry)
    auto write_out = [&](int bw) {
        const bool wact = bw >= 4 * R && bw <= K - 1 - 4 * R;
        const int mub = r0 + BR * bw;
        const int gb = (mub + 3) * P * (int)sizeof(T);
        const T *wbase = ring + slot_of(bw) * BR * W;
#pragma unroll
        for (int j = 0; j < NSV; ++j) {
            const int rr = cinfo[j] & 7;
            const int mu = mub + rr;
            const bool ok = wact && cl[j] >= 0 && mu >= a && mu < aend;
            u32x4 v = *reinterpret_cast<const u32x4 *>(wbase + (cl[j] >= 0 ? cl[j] : 0));
            if ((cinfo[j] >> 3) != 63) {
                Chunk<T> c = __builtin_bit_cast(Chunk<T>, v);
                const T *row = wbase + rr * W;
                const int e1 = (cinfo[j] >> 3) & 7, e2 = (cinfo[j] >> 6) & 7;
                const int c0 = nni + 3;
                const T v1 = even ? d.ex[0] * row[5] + d.ex[1] * row[6] + d.ex[2] * row[7] + d.ex[3] * row[8] +
                                        d.ex[4] * row[9]
                                  : T(0);
                const T v2 = even ? d.ex[0] * row[c0 - 1] + d.ex[1] * row[c0 - 2] + d.ex[2] * row[c0 - 3] +
                                        d.ex[3] * row[c0 - 4] + d.ex[4] * row[c0 - 5]
                                  : T(0);
#pragma unroll
                for (int k = 0; k < VEC; ++k) c.v[k] = k == e1 ? v1 : (k == e2 ? v2 : c.v[k]);
                v = __builtin_bit_cast(u32x4, c);
            }
            __builtin_amdgcn_raw_buffer_store_b128(v, rs_dst, oob_unless(ok, gb + (tid + j * nth) * 16), 0, 0);
        }
    };

    // prologue: coefficients, d(nu), blocks -R .. R and their row weights
    if (tid < 8) coef_s[tid] = d.c[8 + tid];
    if (tid == 8) coef_s[8] = d.omega1;
    for (int e = tid; e < nni; e += nth) dnu_s[e] = d.dnu[e];
    for (int b = -R; b <= R; ++b) {
        u32x4 pr[NSV];
        T bv;
        load_block(b, pr, bv);
        commit_block(b, pr, bv);
    }

    // Lane roles: pass waves (wave < 5, pass = wave, uniform) and spill lanes
    // q (pass q / MSp, point 64 NT + q % MSp for q < 5 MSp, else pass 5 =
    // idle; MSp = this problem's points per pass and row beyond 64 NT, at
    // most the launch's MS: a problem narrower than the launch's widest
    // leaves whole spill waves idle).  Per point j:
    // tq = t clamped to M-1 (keeps the LDS reads of absent points inside the
    // row), lim = 5t for a point of the grid's colour blocks, else a value no
    // row accepts, and toff = its byte offset in a row of (omega/lhs, rhs)
    // pairs, or kOOBLane for an absent point.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool pass_wave = wave < 5;
    const int q = tid - 320;
    const int MSp = max(0, min(MS, M - 64 * NT));
    const int p_spill = pass_wave || q >= 5 * MSp ? 5 : q / MSp;
    constexpr int kNoPoint = 1 << 20;
    int tq[NT], lim[NT];
    unsigned toff[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int t = pass_wave ? (tid & 63) + 64 * j : (j == 0 && p_spill < 5 ? 64 * NT + q % MSp : M);
        const bool ex = t < M;
        tq[j] = min(t, M - 1);
        lim[j] = ex ? 5 * t : kNoPoint;
        toff[j] = ex ? (unsigned)(t * (int)(2 * sizeof(T))) : kOOBLane;
    }
    const int M5 = 5 * M;
    // A spill wave none of whose lanes holds a point of this problem (a
    // problem narrower than the launch's widest) skips the compute phase;
    // its loads and stores stay, so every wave issues the same memory
    // operations.
    const bool spill_idle = !pass_wave && __ballot(lim[0] != kNoPoint) == 0;

    // (omega/lhs, rhs) of pass p's points in step s (rows of block
    // s - SKEW p); rows outside the pass's range or the grid and absent points
    // read zeros, which are never used
    auto load_lr = [&](int s, int p, T (&lv)[BR][NT], T (&rv)[BR][NT]) {
        const int b = s - SKEW * p;
        const bool act = p < 5 && b >= R * p && b <= K - 1 - R * p;
        const int mu0 = r0 + BR * b;
#pragma unroll
        for (int rr = 0; rr < BR; ++rr) {
            const int mu = mu0 + rr;
            const bool rowok = act && mu >= 2 && mu <= nmu - 4;
            const unsigned rb = rowok ? (unsigned)(((mu - 2) * M5 + p * M) * (int)(2 * sizeof(T))) : kOOB;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const Pair<T> pr = buf_load2(rs_lr, rb + toff[j], T(0));
                lv[rr][j] = pr.a;
                rv[rr][j] = pr.b;
            }
        }
    };

    __syncthreads();  // the prologue's LDS contents are in place
    // per-lane copies of the stencil coefficients (read from LDS, so they
    // live in vector registers rather than being re-fetched as scalars)
    T cf[8];  // c[8..15]: d/dnu (c[8..11]) and d2/dnu2 (c[12..15]) weights
#pragma unroll
    for (int k = 0; k < 8; ++k) cf[k] = coef_s[k];
    const T om1 = coef_s[8];

    // pass p on block s - SKEW p with the (omega/lhs, rhs) of step s; within an
    // active pass the points are processed without branches (points outside
    // the grid compute on in-bounds values and write to the sink).  The points
    // of one pass in a row do not read each other (a point's neighbours have
    // other colours), so a row's stencil values are gathered first and its
    // arithmetic follows, which keeps the LDS reads in flight together.
    auto compute_rows = [&](int b, int p, const T (&lv)[BR][NT], const T (&rv)[BR][NT], auto contig) {
        constexpr bool CONTIG = decltype(contig)::value;
        const int sb = slot_of(b);
        const T *const wb = ring + slot_of(b - R) * BR * W;  // the window's first block (CONTIG)
        // the ring blocks b-R .. b+R
        const T *blk[2 * R + 1];
#pragma unroll
        for (int k = -R; k <= R; ++k) {
            int sl = sb + k;
            sl = sl < 0 ? sl + RB : (sl >= RB ? sl - RB : sl);
            blk[k + R] = ring + sl * BR * W;
        }
        T *const b0 = ring + sb * BR * W;         // block b
        const int mu0 = r0 + BR * b;
        const T *ab = ab_r + sb * 8 * BR;        // row weights of block b
        int off = mod5(p - cA * (mu0 - 2));     // (nu-2) mod 5 of the pass's points in row mu0
        // Rows are processed in groups of G: the stencil values of a group are
        // gathered before its arithmetic (the points of one pass never read
        // each other, in one row or in neighbouring rows).
        constexpr int G = 2;
#pragma unroll
        for (int rg = 0; rg < BR; rg += G) {
            int cbv[G][NT], ov[G][NT];
            bool okv[G][NT];
            T x[G][NT][17];  // 0..8: own row at nu-4..nu+4; 9..12: rows mu-1..mu-4; 13..16: rows mu+1..mu+4
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int rr = rg + g;
                if (rr) off = off - cA < 0 ? off - cA + 5 : off - cA;
                const int mu = mu0 + rr;
                const int lim_r = mu >= 2 && mu <= nmu - 4 ? nni - 3 - off : -1;  // point ok iff 5t <= lim_r
#pragma unroll
                for (int j = 0; j < NT; ++j) {
                    const int cb = 1 + off + 5 * tq[j];  // = nu - 1 = (column of nu) - 4
                    cbv[g][j] = cb;
                    okv[g][j] = lim[j] <= lim_r;
                    const T *q0 = b0 + cb;
                    if constexpr (CONTIG) {
                        // all 17 values from one opaque base at the window's
                        // first row, with immediate offsets (own row 4 R ... of
                        // the window: R BR = 4 rows above block b)
                        int o = (int)(wb - dnu_s) + cb;  // from the start of the LDS
                        asm("" : "+v"(o));
                        ov[g][j] = o + (4 + rr) * W;
#pragma unroll
                        for (int k = 0; k < 9; ++k) x[g][j][k] = dnu_s[o + (4 + rr) * W + k];
#pragma unroll
                        for (int k = 1; k <= 4; ++k) {
                            x[g][j][8 + k] = dnu_s[o + (4 + rr - k) * W + 4];
                            x[g][j][12 + k] = dnu_s[o + (4 + rr + k) * W + 4];
                        }
                        continue;
                    } else if constexpr (sizeof(T) == 4) {
                        // the own row's nine values from one opaque base, so that
                        // they are read with immediate offsets of that base
                        int o = (int)(b0 - dnu_s) + rr * W + cb;  // from the start of the LDS
                        asm("" : "+v"(o));
                        ov[g][j] = o;
#pragma unroll
                        for (int k = 0; k < 9; ++k) x[g][j][k] = dnu_s[o + k];
                    } else {
#pragma unroll
                        for (int k = 0; k < 9; ++k) x[g][j][k] = q0[rr * W + k];
                    }
#pragma unroll
                    for (int k = 1; k <= 4; ++k) {
                        // rows rr -/+ k: block offset and row within it (compile-time)
                        const int rdn = rr - k, rup = rr + k;
                        const int bdn = rdn >= 0 ? 0 : -((-rdn + BR - 1) / BR), bup = rup / BR;
                        x[g][j][8 + k] = blk[bdn + R][cb + (rdn - bdn * BR) * W + 4];
                        x[g][j][12 + k] = blk[bup + R][cb + (rup - bup * BR) * W + 4];
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int rr = rg + g;
                const T am4 = ab[8 * rr + 0], am3 = ab[8 * rr + 1], am2 = ab[8 * rr + 2], am1 = ab[8 * rr + 3];
                const T ap4 = ab[8 * rr + 4], ap3 = ab[8 * rr + 5], ap2 = ab[8 * rr + 6], ap1 = ab[8 * rr + 7];
#pragma unroll
                for (int j = 0; j < NT; ++j) {
                    const T *xv = x[g][j];
                    const int cb = cbv[g][j];
                    if constexpr (sizeof(T) == 4) {
                        // one accumulator: the 8 mu terms, the 4 symmetric nu terms,
                        // then d(nu) times the antisymmetric nu sum (28 operations)
                        float acc = am4 * xv[12];
                        acc = __builtin_fmaf(ap4, xv[16], acc);
                        acc = __builtin_fmaf(am3, xv[11], acc);
                        acc = __builtin_fmaf(ap3, xv[15], acc);
                        acc = __builtin_fmaf(am2, xv[10], acc);
                        acc = __builtin_fmaf(ap2, xv[14], acc);
                        acc = __builtin_fmaf(am1, xv[9], acc);
                        acc = __builtin_fmaf(ap1, xv[13], acc);
                        acc = __builtin_fmaf(cf[4], xv[0] + xv[8], acc);
                        acc = __builtin_fmaf(cf[5], xv[1] + xv[7], acc);
                        acc = __builtin_fmaf(cf[6], xv[2] + xv[6], acc);
                        acc = __builtin_fmaf(cf[7], xv[3] + xv[5], acc);
                        float t = cf[0] * (xv[0] - xv[8]);
                        t = __builtin_fmaf(cf[1], xv[1] - xv[7], t);
                        t = __builtin_fmaf(cf[2], xv[2] - xv[6], t);
                        t = __builtin_fmaf(cf[3], xv[3] - xv[5], t);
                        acc = __builtin_fmaf(dnu_s[cb], t, acc);
                        const float v = __builtin_fmaf(rv[rr][j] - acc, lv[rr][j], om1 * xv[4]);
                        T *w = okv[g][j] ? dnu_s + ov[g][j] + 4 : sink + tid;
                        *w = v;
                    } else {
                        const T dmu = (am4 * xv[12] + ap4 * xv[16]) + (am3 * xv[11] + ap3 * xv[15]) +
                                      ((am2 * xv[10] + ap2 * xv[14]) + (am1 * xv[9] + ap1 * xv[13]));
                        const T ddni2 = (cf[4] * (xv[0] + xv[8]) + cf[5] * (xv[1] + xv[7])) +
                                        (cf[6] * (xv[2] + xv[6]) + cf[7] * (xv[3] + xv[5]));
                        const T ddni1 = (cf[0] * (xv[0] - xv[8]) + cf[1] * (xv[1] - xv[7])) +
                                        (cf[2] * (xv[2] - xv[6]) + cf[3] * (xv[3] - xv[5]));
                        const T cc = (dmu + ddni2) + dnu_s[cb] * ddni1;
                        const T v = (rv[rr][j] - cc) * lv[rr][j] + om1 * xv[4];
                        T *w = okv[g][j] ? b0 + cb + rr * W + 4 : sink + tid;
                        *w = v;
                    }
                }
            }
        }
    };
    //
    // The window of blocks b-R .. b+R occupies consecutive ring slots except
    // when it wraps around the ring's end.  A pass wave (its block is a
    // scalar) whose window does not wrap reads all 17 stencil values of a
    // point from one base register with immediate offsets (FP32); otherwise
    // each block has its own base.
    auto compute = [&](int s, int p, const T (&lv)[BR][NT], const T (&rv)[BR][NT], auto role) {
        const int b = s - SKEW * p;
        if (p >= 5 || b < R * p || b > K - 1 - R * p) return;  // idle lanes and passes outside their range
        const int sw = slot_of(b - R);
        if constexpr (decltype(role)::value && sizeof(T) == 4) {
            if (sw + 2 * R < RB) {
                compute_rows(b, p, lv, rv, std::true_type{});
                return;
            }
        }
        compute_rows(b, p, lv, rv, std::false_type{});
    };
    // The mu=1 row (sormcsor.F90:333-350) of the first strip, after its last
    // step: rows 2..6 are final in dst, written by this workgroup (the
    // barrier makes its stores visible to all of its threads).  A fixed number
    // of loads and stores per thread (nni < W <= RT * 320 <= RT * nth; lanes
    // beyond nni store out of range).
    auto row_one = [&]() {
        constexpr int RT = (W + 319) / 320;
        auto r1 = [&](int nu) -> T {
            T acc = T(0);
#pragma unroll
            for (int k = 0; k < 5; ++k)
                acc += d.ex[k] * buf_load1(rs_dst, (unsigned)(((5 + k) * P + nu + 3) * (int)sizeof(T)), T(0));
            return acc;
        };
#pragma unroll
        for (int k = 0; k < RT; ++k) {
            const int nu0 = 1 + tid + k * nth;
            const bool ok = nu0 <= nni;
            const int nu = ok ? nu0 : 2;
            T v = T(0);
            if (even) {
                if (nu == 1 || nu == nni) {
                    const int dir = nu == 1 ? 1 : -1;
                    v = d.ex[0] * r1(nu + dir) + d.ex[1] * r1(nu + 2 * dir) + d.ex[2] * r1(nu + 3 * dir) +
                        d.ex[3] * r1(nu + 4 * dir) + d.ex[4] * r1(nu + 5 * dir);
                } else {
                    v = r1(nu);
                }
            }
            buf_store1(rs_dst, oob_unless(ok, (4 * P + nu + 3) * (int)sizeof(T)), v);
        }
    };

    // Block s+R+1 and the (omega/lhs, rhs) of step s+1 are loaded at the
    // start of step s and land during its compute phase (the compiler's
    // waits hold back on them at their first use, after the compute).  Block
    // bw = s - 4 SKEW - R is final from the start of step s (pass 4 wrote it
    // in step s-R; in step s it only reads it, as a neighbour of block bw+R).
    // FP64 writes it out right after issuing the loads, so its stores overlap
    // the compute phase; FP32 writes it out at the end of the step (each
    // measured faster that way).  The commit of block s+R+1 reuses the slot of
    // block s+R+1-RB = s-4 SKEW-R-1, written out in step s-1.
    constexpr bool early_out = sizeof(T) == 8;
    // The two lane roles run separate copies of the step loop (each wave is
    // wholly in one role and every copy meets the same barriers), so that in
    // a pass wave the pass and all that depends on it stay scalar.
    auto steps = [&](auto role) {
        constexpr bool PASS = decltype(role)::value;
        const int p = PASS ? wave : p_spill;
        const bool busy = PASS || !spill_idle;
        T lv[BR][NT], rv[BR][NT];
        load_lr(0, p, lv, rv);
        for (int s = 0; s < S; ++s) {
            T ln[BR][NT], rn[BR][NT], bpre;
            u32x4 pre[NSV];
            load_lr(s + 1, p, ln, rn);
            load_block(s + R + 1, pre, bpre);
            if (early_out) write_out(s - 4 * SKEW - R);
            if (busy) compute(s, p, lv, rv, role);
            commit_block(s + R + 1, pre, bpre);
#pragma unroll
            for (int rr = 0; rr < BR; ++rr)
#pragma unroll
                for (int j = 0; j < NT; ++j) {
                    lv[rr][j] = ln[rr][j];
                    rv[rr][j] = rn[rr][j];
                }
            if (!early_out) write_out(s - 4 * SKEW - R);
            __syncthreads();
        }
    };
    if (pass_wave) steps(std::true_type{});
    else steps(std::false_type{});
    if (a == 2) {
        __syncthreads();
        row_one();
    }
}

// ---------------------------------------------------------------------------
// putin / putout (inout.F90:23-82, 165-183), several problems per launch.
struct PadDesc {
    const double *fun;
    double *w1;
    double *w2;  // optional second copy (ping-pong partner), may be null
    int nni, nmu, isym;
    int ps;      // row stride of w1 / w2 (>= nni+8; columns beyond nni+8 are zero)
};

__global__ void k_putin(const PadDesc *__restrict__ pd)
{
    const PadDesc &d = pd[blockIdx.y];
    const int nni = d.nni, nmu = d.nmu, PS = d.ps;
    const long tot = (long)PS * (nmu + 8);
    const double s = d.isym == 1 ? 1.0 : -1.0;
    // 32-bit index arithmetic (a padded grid has fewer than 2^31 points)
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < (int)tot; e += gridDim.x * blockDim.x) {
        const int P = e / PS, Q = e - P * PS;
        const int mu = P - 3, nu = Q - 3;   // 1-based grid indices, halo range -3..n+4
        double v = 0.0;
        const bool nu_in = nu >= 1 && nu <= nni, mu_in = mu >= 1 && mu <= nmu;
        if (nu_in && mu_in) {
            v = d.fun[(long)(mu - 1) * nni + nu - 1];
        } else if (nu_in && mu >= -3 && mu <= 0) {
            v = s * d.fun[(long)(1 - mu) * nni + nu - 1];          // mu' = 2 - mu
        } else if (mu_in && nu >= -3 && nu <= 0) {
            v = s * d.fun[(long)(mu - 1) * nni + (1 - nu)];        // nu' = 2 - nu
        } else if (mu_in && nu >= nni + 1 && nu <= nni + 4) {
            v = s * d.fun[(long)(mu - 1) * nni + (2 * nni - nu - 1)];  // nu' = 2nni - nu
        }
        d.w1[e] = v;
        if (d.w2) d.w2[e] = v;
    }
}

__global__ void k_putout(const PadDesc *__restrict__ pd)
{
    const PadDesc &d = pd[blockIdx.y];
    const int nni = d.nni, nmu = d.nmu, PS = d.ps;
    const long tot = (long)nni * nmu;
    double *fun = const_cast<double *>(d.fun);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
         e += (long)gridDim.x * blockDim.x) {
        const int mu = 1 + (int)(e / nni), nu = 1 + (int)(e % nni);
        fun[e] = d.w1[(long)(mu + 3) * PS + nu + 3];
    }
}

// A padded array copied between row strides (destination columns beyond
// the source's are zero), into one or two destinations.
struct RestrideDesc {
    const double *src;
    double *dst1, *dst2;  // dst2 may be null
    int cols, rows;       // nni+8, nmu+8
    int ss, ds;           // source / destination row strides (ds >= cols)
};

__global__ void k_restride(const RestrideDesc *__restrict__ rd)
{
    const RestrideDesc &d = rd[blockIdx.y];
    const long tot = (long)d.ds * d.rows;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e / d.ds), c = (int)(e - (long)r * d.ds);
        const double v = c < d.cols ? d.src[(long)r * d.ss + c] : 0.0;
        d.dst1[e] = v;
        if (d.dst2) d.dst2[e] = v;
    }
}

// ---------------------------------------------------------------------------
// Orbital tail (orbAsymptGet / orbAsymptSet, relaxOrbs.F90:104-124, 415-462):
// edecay from fa = fock1/f1 (zero where |f1| <= eps), then
// psi(nu,mu) = psi(nu,mu-1)*edecay for mu = nmu-3..nmu, sequentially in mu.
struct TailDesc {
    double *psi;
    const double *fock1;
    const double *f1;
    const double *vxisq;
    const double *vetasq;
    int nni, nmu;
};

__global__ void k_orb_tail(const TailDesc *__restrict__ td)
{
    const TailDesc &d = td[blockIdx.y];
    const int nni = d.nni, nmu = d.nmu;
    const double eps = 2.220446049250313e-16;  // epsilon(1.0d0), setPrecision
    for (int nu = blockIdx.x * blockDim.x + threadIdx.x + 1; nu <= nni; nu += gridDim.x * blockDim.x) {
        double fa[5];
        for (int k = 0; k < 5; ++k) {
            const int mu = nmu - 4 + k;
            const long u = (long)(mu - 1) * nni + nu - 1;
            const double f1 = d.f1[u];
            fa[k] = fabs(f1) <= eps ? 0.0 : d.fock1[u] / f1;
        }
        double ed[4];
        for (int jj = 0; jj < 4; ++jj) {
            const int mu = nmu - 3 + jj;
            const double raiq1 = sqrt(d.vxisq[mu - 2] + d.vetasq[nu - 1] - 1.0);
            const double raiq = sqrt(d.vxisq[mu - 1] + d.vetasq[nu - 1] - 1.0);
            ed[jj] = pow(fabs(fa[jj + 1] / fa[jj]), 0.25) * exp(sqrt(fabs(fa[jj])) * (raiq1 - raiq));
        }
        for (int jj = 0; jj < 4; ++jj) {
            const int mu = nmu - 3 + jj;
            d.psi[(long)(mu - 1) * nni + nu - 1] = d.psi[(long)(mu - 2) * nni + nu - 1] * ed[jj];
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed precision helpers.
// u(nu, mu) of an unpadded function with the mirror halos of putin applied on
// the fly (inout.F90:51-63): below mu = 1 and outside nu = 1..nni the value
// is isym times the mirrored one.  The 17-point cross needs at most one mirror.
__device__ __forceinline__ double fmirror(const double *f, int nu, int mu, int nni, double s)
{
    double sc = 1.0;
    if (mu < 1) {
        mu = 2 - mu;
        sc = s;
    } else if (nu < 1) {
        nu = 2 - nu;
        sc = s;
    } else if (nu > nni) {
        nu = 2 * nni - nu;
        sc = s;
    }
    return sc * f[(long)(mu - 1) * nni + nu - 1];
}

// Residual r = rhs - (L + lhs) u over the relaxed interior (FP64, read from
// the unpadded u with its mirror halos), stored in FP32 for the correction
// sweeps; the operator's terms in the order of sormcsor.F90:173-194.
struct ResDesc {
    const double *fun;     // unpadded u
    const double *lhs;
    const double *rhs;
    const double *rf[3];   // rf[0] != null: rhs = rf[0] * rf[1] * rf[2] (null factors count as 1)
    const double *bmu;
    const double *dnu;
    float *res_f;          // FP32 residual (may be null)
    int res_split;         // res_f holds colour-ordered (omega/lhs, r) pairs (else r, unpadded nni*nmu)
    double omega;          // for the pairs
    double *part;          // per-wave partial sums (res^2, rhs^2) of this problem at
                           // [2 * (block of the problem * waves per block + wave)] (may be null)
    double c[16];
    int nni, nmu, isym;
};

// Workgroup index remapped so that each of the 8 XCDs (which take the
// workgroups of a launch round-robin, each with its own L2) gets one
// contiguous range of the launch's blocks: a point's neighbour rows are then
// read by workgroups of the same XCD and come from its L2.
__device__ __forceinline__ int xcd_block(int b, int total)
{
    const int x = b & 7, slot = b >> 3, per = total >> 3, rem = total & 7;
    return x * per + min(x, rem) + slot;
}

// The residual is formed with the second differences taken on u - u(p):
//   L u + lhs u = sum_k c2_k ((u_-k - u0) + (u_+k - u0)) [mu and nu]
//                 + b d1 u + d d1 u + (lhs + 2 sum_k c2_k) u0,
// the same operator (the 9-point second-difference weights sum to minus the
// diagonal built into lhs), but the large terms c2_k u ~ u / h^2 that cancel
// against lhs u are no longer rounded: neighbour differences of a smooth
// function are exact (Sterbenz) and small, and lhs + 2 sum c2_k is a small
// exact difference.  The FP64 rounding of the residual then stays well
// below the rounding floor of u itself.  Returns 2 sum_k c2_k over mu and nu.
__device__ __forceinline__ double residual_diag_shift(const double *c)
{
    return 2.0 * ((c[4] + c[5]) + (c[6] + c[7])) + 2.0 * ((c[12] + c[13]) + (c[14] + c[15]));
}

// One point per thread; blocks_per_problem blocks of 256 points per problem
// (a 1-D grid of blocks_per_problem * problems workgroups).
__global__ void __launch_bounds__(256) k_residual(const ResDesc *__restrict__ rd, int blocks_per_problem)
{
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const ResDesc &d = rd[lb / blocks_per_problem];
    const int nni = d.nni, nmu = d.nmu;
    const long e = (long)(lb % blocks_per_problem) * 256 + threadIdx.x;
    double sr = 0.0, sb = 0.0;
    if (e < (long)nni * nmu) {
        const int mu = 1 + (int)(e / nni), nu = 1 + (int)(e - (long)(mu - 1) * nni);
        if (mu >= 2 && mu <= nmu - 4 && nu >= 2 && nu <= nni - 1) {
            const double s = d.isym == 1 ? 1.0 : -1.0;
            const double *f = d.fun, *c = d.c;
            double cc;
            const double bm = d.bmu[mu - 1], dn = d.dnu[nu - 1], u0 = f[e];
            // second differences on u - u0, as in k_residual_rows
            if (nu >= 5 && nu <= nni - 4 && mu >= 5) {
                // no mirror images: neighbours straight from u
                const double *q = f + e;
                const long s1 = nni, s2 = 2L * nni, s3 = 3L * nni, s4 = 4L * nni;
                const double ddmi2 = c[4] * ((q[-s4] - u0) + (q[s4] - u0)) + c[5] * ((q[-s3] - u0) + (q[s3] - u0)) +
                                     c[6] * ((q[-s2] - u0) + (q[s2] - u0)) + c[7] * ((q[-s1] - u0) + (q[s1] - u0));
                const double ddmi1 = c[0] * (q[-s4] - q[s4]) + c[1] * (q[-s3] - q[s3]) + c[2] * (q[-s2] - q[s2]) +
                                     c[3] * (q[-s1] - q[s1]);
                const double ddni2 = c[12] * ((q[-4] - u0) + (q[4] - u0)) + c[13] * ((q[-3] - u0) + (q[3] - u0)) +
                                     c[14] * ((q[-2] - u0) + (q[2] - u0)) + c[15] * ((q[-1] - u0) + (q[1] - u0));
                const double ddni1 = c[8] * (q[-4] - q[4]) + c[9] * (q[-3] - q[3]) + c[10] * (q[-2] - q[2]) +
                                     c[11] * (q[-1] - q[1]);
                cc = ddmi2 + bm * ddmi1 + ddni2 + dn * ddni1;
            } else {
                auto F = [&](int dnu, int dmu) { return fmirror(f, nu + dnu, mu + dmu, nni, s) - u0; };
                auto G = [&](int dnu, int dmu) { return fmirror(f, nu + dnu, mu + dmu, nni, s); };
                const double ddmi2 = c[4] * (F(0, -4) + F(0, 4)) + c[5] * (F(0, -3) + F(0, 3)) +
                                     c[6] * (F(0, -2) + F(0, 2)) + c[7] * (F(0, -1) + F(0, 1));
                const double ddmi1 = c[0] * (G(0, -4) - G(0, 4)) + c[1] * (G(0, -3) - G(0, 3)) +
                                     c[2] * (G(0, -2) - G(0, 2)) + c[3] * (G(0, -1) - G(0, 1));
                const double ddni2 = c[12] * (F(-4, 0) + F(4, 0)) + c[13] * (F(-3, 0) + F(3, 0)) +
                                     c[14] * (F(-2, 0) + F(2, 0)) + c[15] * (F(-1, 0) + F(1, 0));
                const double ddni1 = c[8] * (G(-4, 0) - G(4, 0)) + c[9] * (G(-3, 0) - G(3, 0)) +
                                     c[10] * (G(-2, 0) - G(2, 0)) + c[11] * (G(-1, 0) - G(1, 0));
                cc = ddmi2 + bm * ddmi1 + ddni2 + dn * ddni1;
            }
            const double l = d.lhs[e];
            double b;
            if (d.rf[0]) {
                b = d.rf[0][e];
                if (d.rf[1]) b *= d.rf[1][e];
                if (d.rf[2]) b *= d.rf[2][e];
            } else {
                b = d.rhs[e];
            }
            const double r = b - (cc + (l + residual_diag_shift(c)) * u0);
            sr = r * r;
            sb = b * b;
            if (d.res_f) {
                if (d.res_split) {
                    float2 pr;
                    pr.x = (float)(d.omega / l);
                    pr.y = (float)r;
                    reinterpret_cast<float2 *>(d.res_f)[split_index(nu, mu, split_m(nni), colour_mult(nni))] = pr;
                } else {
                    d.res_f[e] = (float)r;
                }
            }
        }
    }
    if (d.part) {
        // fixed-order wavefront reduction; each wave owns one partial slot
        for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_down(sr, off, 64);
            sb += __shfl_down(sb, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            const int slot = (lb % blocks_per_problem) * (blockDim.x >> 6) + (threadIdx.x >> 6);
            d.part[2 * slot] = sr;
            d.part[2 * slot + 1] = sb;
        }
    }
}

// The same residual with u read once: a workgroup owns the nu columns of a
// strip of kResRows relaxed rows of one problem (thread t <-> nu = t+1), keeps
// u(nu, mu-4..mu+4) in registers as it walks down the strip, and reads the nu
// neighbours from the current row staged in LDS (two row buffers, one barrier
// per row).  The mirror images of the mu < 1 rows and of the nu < 1, nu > nni
// columns are placed as they are read, so every point sees the values of the
// per-point kernel and combines them in the same order.  Needs nni <= blockDim.
//
// The loads of the next kResAhead rows are in flight while a group of rows
// is computed.  Every lane issues every load (rows clamped to the strip,
// columns to the grid; buffer loads, which only the vector-memory counter
// tracks) and every store (through a buffer resource, at an out-of-range
// offset for lanes without a relaxed point), so each wave issues the same
// memory operations in every row and the compiler's waits count exactly.
// One row ahead and 32-row strips measured fastest: deeper prefetch costs
// more in occupancy (registers) than it gains in latency hiding.
constexpr int kResRows = 32;
constexpr int kResThreads = 512;
constexpr int kResAhead = 1;

// PROD: some problem of the launch gives its rhs as a product of factors
// (those that do not read rhs alone, with the same operation count).
template <bool PROD>
__global__ void __launch_bounds__(kResThreads) k_residual_rows(const ResDesc *__restrict__ rd, int strips)
{
    __shared__ double row[2][kResThreads + 8];
    const ResDesc &d = rd[blockIdx.x / strips];
    const int nni = d.nni, nmu = d.nmu;
    const int m0 = 2 + (blockIdx.x % strips) * kResRows;
    if (m0 > nmu - 4) return;                      // uniform over the workgroup
    const int m1 = min(m0 + kResRows - 1, nmu - 4);
    const int nu = threadIdx.x + 1;
    const bool col = nu <= nni;
    const bool interior = nu >= 2 && nu <= nni - 1;
    const int cl = min(nu, nni) - 1;               // column read by this lane
    const double s = d.isym == 1 ? 1.0 : -1.0;
    const double *c = d.c;
    const double shift = residual_diag_shift(c);
    // u, lhs and rhs through buffer resources: vector-memory loads, waited
    // for by their own counter only
    const unsigned nbytes = (unsigned)((long)nni * nmu * 8);
    // the right-hand side is rhs, or the product of up to three factors
    // (loads of absent factors fall outside a zero-size buffer; unused)
    const bool prod = PROD && d.rf[0] != nullptr, has1 = prod && d.rf[1], has2 = prod && d.rf[2];
    const __amdgpu_buffer_rsrc_t rs_u = buf_rsrc(d.fun, nbytes), rs_l = buf_rsrc(d.lhs, nbytes),
                                 rs_b = buf_rsrc(prod ? d.rf[0] : d.rhs, nbytes),
                                 rs_b1 = buf_rsrc(d.rf[1], has1 ? nbytes : 0u),
                                 rs_b2 = buf_rsrc(d.rf[2], has2 ? nbytes : 0u);
    auto load_u = [&](int mu) -> double {         // u(nu, mu) for -3 <= mu <= nmu, mirrored below mu = 1
        const double v = buf_load1(rs_u, (unsigned)(((mu < 1 ? 1 - mu : mu - 1) * nni + cl) * 8), 0.0);
        return mu < 1 ? s * v : v;
    };
    const int M = split_m(nni), cA = colour_mult(nni);
    const bool out = d.res_f != nullptr;
    const __amdgpu_buffer_rsrc_t rs_out =
        buf_rsrc(d.res_f, out ? (unsigned)(d.res_split ? split_size(nni, nmu) * 8 : (long)nni * nmu * 4) : 0u);
    double w[8];                                   // u(nu, mu-4 .. mu+3)
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = load_u(m0 - 4 + k);
    const __amdgpu_buffer_rsrc_t rs_bm = buf_rsrc(d.bmu, (unsigned)(nmu * 8));
    const double dn = buf_load1(buf_rsrc(d.dnu, (unsigned)(nni * 8)), (unsigned)(cl * 8), 0.0);
    double sr = 0.0, sb = 0.0;
    // the rows are taken in groups of kResAhead: while a group is computed,
    // the next group's u(nu, mu+4), lhs(nu, mu), rhs(nu, mu) and b(mu) are in
    // flight (rows past m1 read m1); the group's values are then moved into
    // place once all have arrived, so no load in flight is ever moved
    double up[kResAhead], lp[kResAhead], bp[kResAhead], mp[kResAhead];
    auto issue = [&](double (&u_)[kResAhead], double (&l_)[kResAhead], double (&b_)[kResAhead],
                     double (&m_)[kResAhead], int mb) {
#pragma unroll
        for (int k = 0; k < kResAhead; ++k) {
            const int mr = min(mb + k, m1);
            const unsigned e8 = (unsigned)(((mr - 1) * nni + cl) * 8);
            u_[k] = load_u(mr + 4);
            l_[k] = buf_load1(rs_l, e8, 0.0);
            b_[k] = buf_load1(rs_b, e8, 0.0);
            if constexpr (PROD) {
                // issued by every problem of the launch (same operation count
                // in every row)
                const double f1 = buf_load1(rs_b1, e8, 0.0), f2 = buf_load1(rs_b2, e8, 0.0);
                b_[k] = has1 ? b_[k] * f1 : b_[k];
                b_[k] = has2 ? b_[k] * f2 : b_[k];
            }
            m_[k] = buf_load1(rs_bm, (unsigned)((mr - 1) * 8), 0.0);
        }
    };
    issue(up, lp, bp, mp, m0);
    // rows past m1 of the last group run with clamped loads and neither
    // store nor count (no early exit: every wave issues the same operations)
    for (int mb = m0; mb <= m1; mb += kResAhead) {
        double un[kResAhead], ln[kResAhead], bn[kResAhead], mn[kResAhead];
        issue(un, ln, bn, mn, mb + kResAhead);
#pragma unroll
        for (int k = 0; k < kResAhead; ++k) {
            const int mu = mb + k;
            const bool live = interior && mu <= m1;
            // stage row mu with its mirrored columns: slot nu+3 for nu = -3 .. nni+4
            double *r = row[mu & 1];
            if (col) {
                r[nu + 3] = w[4];
                if (nu >= 2 && nu <= 5) r[5 - nu] = s * w[4];                        // nu' = 2 - nu
                if (nu >= nni - 4 && nu <= nni - 1) r[2 * nni - nu + 3] = s * w[4];  // nu' = 2 nni - nu
            }
            __syncthreads();
            // every lane computes (slots nu-1 .. nu+7 lie inside the row
            // buffer); only the relaxed points of rows <= m1 count and store
            const double *q = r + nu + 3;
            const double w8 = up[k], l = lp[k], b = bp[k], u0 = w[4];
            // second differences on u - u0 (see residual_diag_shift)
            const double ddmi2 = c[4] * ((w[0] - u0) + (w8 - u0)) + c[5] * ((w[1] - u0) + (w[7] - u0)) +
                                 c[6] * ((w[2] - u0) + (w[6] - u0)) + c[7] * ((w[3] - u0) + (w[5] - u0));
            const double ddmi1 = c[0] * (w[0] - w8) + c[1] * (w[1] - w[7]) + c[2] * (w[2] - w[6]) +
                                 c[3] * (w[3] - w[5]);
            const double ddni2 = c[12] * ((q[-4] - u0) + (q[4] - u0)) + c[13] * ((q[-3] - u0) + (q[3] - u0)) +
                                 c[14] * ((q[-2] - u0) + (q[2] - u0)) + c[15] * ((q[-1] - u0) + (q[1] - u0));
            const double ddni1 = c[8] * (q[-4] - q[4]) + c[9] * (q[-3] - q[3]) + c[10] * (q[-2] - q[2]) +
                                 c[11] * (q[-1] - q[1]);
            const double cc = ddmi2 + mp[k] * ddmi1 + ddni2 + dn * ddni1;
            const double res = b - (cc + (l + shift) * u0);
            sr += live ? res * res : 0.0;
            sb += live ? b * b : 0.0;
            if (out) {                             // uniform over the launch
                if (d.res_split) {
                    const long o = live ? split_index(nu, mu, M, cA) * 8 : 0;
                    const float2 pr = make_float2((float)(d.omega / l), (float)res);
                    buf_store1(rs_out, oob_unless(live, (int)o), __builtin_bit_cast(double, pr));
                } else {
                    const long o = ((long)(mu - 1) * nni + nu - 1) * 4;
                    buf_store1(rs_out, oob_unless(live, (int)o), (float)res);
                }
            }
#pragma unroll
            for (int k2 = 0; k2 < 7; ++k2) w[k2] = w[k2 + 1];
            w[7] = w8;
        }
#pragma unroll
        for (int k = 0; k < kResAhead; ++k) {
            up[k] = un[k];
            lp[k] = ln[k];
            bp[k] = bn[k];
            mp[k] = mn[k];
        }
    }
    if (d.part) {
        for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_down(sr, off, 64);
            sb += __shfl_down(sb, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            const int slot = (blockIdx.x % strips) * (blockDim.x >> 6) + (threadIdx.x >> 6);
            d.part[2 * slot] = sr;
            d.part[2 * slot + 1] = sb;
        }
    }
}

// out[2 i], out[2 i + 1] = the sums of problem i's nslot partial pairs, in a
// fixed order (one wave per problem: lane-strided sums, then a fixed tree).
__global__ void __launch_bounds__(64) k_sum_parts(const double *__restrict__ part, int nslot, double *__restrict__ out)
{
    const double *p = part + (size_t)2 * nslot * blockIdx.x;
    double sr = 0.0, sb = 0.0;
    for (int k = threadIdx.x; k < nslot; k += 64) {
        sr += p[2 * k];
        sb += p[2 * k + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_down(sr, off, 64);
        sb += __shfl_down(sb, off, 64);
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sr;
        out[2 * blockIdx.x + 1] = sb;
    }
}

// Launch the residual of n problems (descriptors d_res) on grids of at most
// maxnni x maxnmu points: the row kernel when a row fits one workgroup, else
// the per-point kernel; prod: some problem's rhs is a product of factors.
static int launch_residual(x2g_ctx *ctx, const ResDesc *d_res, int n, long maxnni, long maxnmu, long maxn,
                           bool prod)
{
    if (maxnni <= kResThreads) {
        const int strips = (int)((maxnmu - 5 + kResRows - 1) / kResRows);
        const int nth = (int)((maxnni + 63) / 64 * 64);
        if (prod)
            X2G_LAUNCH(ctx, k_residual_rows<true>, dim3(strips * n), dim3(nth), 0, d_res, strips);
        else
            X2G_LAUNCH(ctx, k_residual_rows<false>, dim3(strips * n), dim3(nth), 0, d_res, strips);
    } else {
        const int bpp = (int)((maxn + 255) / 256);
        X2G_LAUNCH(ctx, k_residual, dim3(bpp * n), dim3(256), 0, d_res, bpp);
    }
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

struct CorrDesc {
    double *fun;        // unpadded FP64 u, corrected in place
    const float *e;     // padded FP32 correction
    int nni, nmu;
    int ps;             // row stride of e
};

// u += e over the relaxed points, a row per workgroup at a time (row
// r: mu = 2 + r; column c: nu = 2 + c; 32-bit index arithmetic: a grid has
// fewer than 2^31 points; global loads and stores).
__global__ void __launch_bounds__(256) k_correct(const CorrDesc *__restrict__ cd)
{
    const CorrDesc &d = cd[blockIdx.y];
    const int nni = d.nni, w = nni - 2, s1 = d.ps, rows = d.nmu - 5;
    const gfptr e = gmem(d.e);
    const gdptr_w u = gmem_w(d.fun);
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int ub = (1 + r) * nni + 1, eb = (5 + r) * s1 + 5;
        for (int c = threadIdx.x; c < w; c += blockDim.x) u[ub + c] += (double)e[eb + c];
    }
}

// The boundary lines of an unpadded function from its (corrected) interior,
// as k_boundary does on the padded array (sormcsor.F90:94-129).
struct BndDesc {
    double *fun;
    double ex[5];
    int nni, nmu, isym;
};

__global__ void k_boundary_fun(const BndDesc *__restrict__ bd)
{
    const BndDesc &d = bd[blockIdx.y];
    const int nni = d.nni, nmu = d.nmu;
    double *f = d.fun;
    auto at = [&](int nu, int mu) -> double & { return f[(long)(mu - 1) * nni + nu - 1]; };
    auto row1 = [&](int nu) {
        return d.ex[0] * at(nu, 2) + d.ex[1] * at(nu, 3) + d.ex[2] * at(nu, 4) + d.ex[3] * at(nu, 5) +
               d.ex[4] * at(nu, 6);
    };
    const int nrow1 = nni - 2, ncol = nmu - 4, total = nrow1 + 2 * ncol;
    const bool even = d.isym == 1;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        if (t < nrow1) {
            const int nu = 2 + t;
            at(nu, 1) = even ? row1(nu) : 0.0;
        } else {
            const int q = t - nrow1;
            const int side = q / ncol, mu = 1 + q % ncol;
            const int nu = side == 0 ? 1 : nni, dir = side == 0 ? 1 : -1;
            double v = 0.0;
            if (even) {
                if (mu == 1) {
                    v = d.ex[0] * row1(nu + dir) + d.ex[1] * row1(nu + 2 * dir) + d.ex[2] * row1(nu + 3 * dir) +
                        d.ex[3] * row1(nu + 4 * dir) + d.ex[4] * row1(nu + 5 * dir);
                } else {
                    v = d.ex[0] * at(nu + dir, mu) + d.ex[1] * at(nu + 2 * dir, mu) +
                        d.ex[2] * at(nu + 3 * dir, mu) + d.ex[3] * at(nu + 4 * dir, mu) +
                        d.ex[4] * at(nu + 5 * dir, mu);
                }
            }
            at(nu, mu) = v;
        }
    }
}

struct CvtDesc {
    const double *src;
    float *dst;
    long n;
};
__global__ void k_to_f32(const CvtDesc *__restrict__ cd)
{
    const CvtDesc &d = cd[blockIdx.y];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (long)gridDim.x * blockDim.x)
        d.dst[e] = (float)d.src[e];
}
struct FillFDesc {
    float *dst;
    long n;
};
__global__ void k_zero_f32(const FillFDesc *__restrict__ fd)
{
    const FillFDesc &d = fd[blockIdx.y];
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < d.n; e += (long)gridDim.x * blockDim.x)
        d.dst[e] = 0.0f;
}

// ---------------------------------------------------------------------------
// Host side.

static inline int grid_x_for(long work, int block) { return (int)std::min<long>((work + block - 1) / block, 1024); }

// Upload a host array of descriptors through the context's staging ring.
template <typename D>
static int stage_descs(x2g_ctx *ctx, const std::vector<D> &v, size_t & /*unused*/, const D **out)
{
    return stage_n(ctx, v.data(), v.size(), out);
}

template <typename T>
struct SweepPlan;
template <typename T>
static int stage_plan(x2g_ctx *ctx, SweepPlan<T> &plan, const SweepDesc<T> **de, const SweepDesc<T> **dodd);

int launch_putin(x2g_grid *g, int isym, const double *fun, double *w1, double *w2)
{
    std::vector<PadDesc> v{{fun, w1, w2, g->nni, g->nmu, isym, g->nni + 8}};
    size_t off = 0;
    const PadDesc *dd;
    int rc = stage_descs(g->ctx, v, off, &dd);
    if (rc) return rc;
    X2G_LAUNCH(g->ctx, k_putin, dim3(grid_x_for((long)g->npad(), 256), 1), dim3(256), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

int launch_putin_items(x2g_ctx *ctx, const std::vector<PadItem> &items)
{
    if (items.empty()) return X2G_OK;
    std::vector<PadDesc> v;
    long maxpad = 0;
    for (auto &it : items) {
        v.push_back({it.fun, it.w, nullptr, it.nni, it.nmu, it.isym, it.nni + 8});
        maxpad = std::max<long>(maxpad, (long)(it.nni + 8) * (it.nmu + 8));
    }
    const PadDesc *dd;
    int rc = stage_n(ctx, v.data(), v.size(), &dd);
    if (rc) return rc;
    X2G_LAUNCH(ctx, k_putin, dim3(grid_x_for(maxpad, 256), (int)v.size()), dim3(256), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

int launch_putout(x2g_grid *g, double *fun, const double *work)
{
    std::vector<PadDesc> v{{fun, const_cast<double *>(work), nullptr, g->nni, g->nmu, 1, g->nni + 8}};
    size_t off = 0;
    const PadDesc *dd;
    int rc = stage_descs(g->ctx, v, off, &dd);
    if (rc) return rc;
    X2G_LAUNCH(g->ctx, k_putout, dim3(grid_x_for((long)g->n(), 256), 1), dim3(256), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

// Sweep-engine selection: 0 = k_sweep where it applies, else k_strip, else
// the per-pass kernels (default); 1 = per-pass kernels; 2 = k_strip where the
// LDS ring fits.  strip_rows > 0 fixes the strip height.  Initial values come
// from X2G_SWEEP=phase|strip and X2G_STRIP_L.
static int g_engine = -1, g_strip_rows = -1;
// k_sweep workgroups per CU for the balanced schedule (0: as many as the LDS
// holds); from X2G_SWEEP_WPC.
static int g_sweep_wpc = 0;
// k_sweep batches launch once per ring-stride class with X2G_SWEEP_SPLIT=1
// (default: one launch in the widest grid's class)
static int g_sweep_split = 0;
// The k_sweep plan's free choices: workgroups per CU the segment planner
// fills (0: as many as the LDS ring allows) and the launch split by ring
// class.  The autotuner picks them per batch shape.
struct SweepKnobs {
    int wpc = 0, split = 0, br = 4;
};
// k_sweep block height (4 or 2 rows; X2G_SWEEP_BR) when not tuned
static int g_sweep_br = 4;

// the plan's choices are fixed (x2g_set_sweep_plan, or any of X2G_SWEEP_BR,
// X2G_SWEEP_WPC, X2G_SWEEP_SPLIT set): the autotuner stays out
static bool g_plan_fixed = false;
static int g_autotune = -1;  // -1: from X2G_AUTOTUNE (default on)
// bumped by every setting above (a captured relaxation graph is only
// replayed under the settings it was captured with)
static std::atomic<long> g_plan_gen{0};
// ordered sweeps: 1 = the pipelined schedule (k_levels_pipe), 0 = one phase
// of one sweep at a time (k_levels; X2G_LEVELS=phase); -1 = not read yet
static int g_levels_pipe = -1;
static bool levels_pipe()
{
    if (g_levels_pipe < 0) {
        const char *e = getenv("X2G_LEVELS");
        g_levels_pipe = !(e && std::strcmp(e, "phase") == 0);
    }
    return g_levels_pipe != 0;
}
static void engine_defaults()
{
    if (g_engine >= 0) return;
    const char *e = getenv("X2G_SWEEP");
    g_engine = 0;
    if (e && std::strcmp(e, "phase") == 0) g_engine = 1;
    if (e && std::strcmp(e, "strip") == 0) g_engine = 2;
    const char *l = getenv("X2G_STRIP_L");
    g_strip_rows = l ? atoi(l) : -1;
    const char *w = getenv("X2G_SWEEP_WPC");
    g_sweep_wpc = w ? atoi(w) : 0;
    const char *brv = getenv("X2G_SWEEP_BR");
    g_sweep_br = brv && atoi(brv) == 2 ? 2 : 4;
    const char *sp = getenv("X2G_SWEEP_SPLIT");
    g_sweep_split = sp ? atoi(sp) != 0 : 0;
    g_plan_fixed = w || brv || sp;
    if (g_autotune < 0) {
        const char *at = getenv("X2G_AUTOTUNE");
        g_autotune = at ? atoi(at) != 0 : 1;
    }
}
static int env_engine()
{
    engine_defaults();
    return g_engine;
}
static int env_strip_l()
{
    engine_defaults();
    return g_strip_rows;
}

template <typename T>
static size_t strip_lds(int nni, int skew) { return (size_t)(4 * skew + 4) * 4 * (nni + 8) * sizeof(T); }

// Choose SKEW (2, 1) or 0 = does not fit the LDS ring.
template <typename T>
static int choose_skew(x2g_ctx *ctx, int nni)
{
    if (4 * (nni + 8) > kMaxLoads * 512) return 0;
    if (strip_lds<T>(nni, 2) <= ctx->max_lds) return 2;
    if (strip_lds<T>(nni, 1) <= ctx->max_lds) return 1;
    return 0;
}

// Strip height for k_strip: enough workgroups to fill the chip, at least 32 rows.
static int choose_L(x2g_ctx *ctx, long total_rows)
{
    const int forced = env_strip_l();
    if (forced > 0) return (forced + 3) & ~3;
    // aim for >= 2 workgroups per CU when the batch allows it
    long target = 2L * ctx->num_cu;
    long L = total_rows / std::max<long>(target, 1);
    L = std::max<long>(32, std::min<long>(128, L));
    return (int)((L + 3) & ~3L);
}

// Row segments of a k_sweep launch.  A segment of L rows costs L/4 + 13
// pipeline steps (its halos and the pipeline fill), whatever problem it
// belongs to, so the launch time is set by the busiest workgroup's sum of
// segment costs.  With a forced strip height every workgroup runs one strip.
// Otherwise the workgroups take consecutive rows of the batch, each up to a
// common budget of B steps, cut into a multiple of 4 rows where a problem
// straddles two workgroups; B is the smallest budget for which the rows fit
// in nwg workgroups (those the chip holds at once: wpc per CU, at most one
// per 8 rows).  Results do not depend on the cut: a strip's rows come out
// the same whatever lies beyond its halos.
static void segments_for_budget(const std::vector<int> &rows, long B, int br, std::vector<SweepSeg> &segs,
                                std::vector<int> &wgseg)
{
    // a segment of L rows costs L/br + kFill steps (the halos and the
    // pipeline fill: 13 with 4-row blocks, 22 with 2-row blocks)
    constexpr int kMinRows = 8;
    const int kFill = 32 / br + 4 * sweep_skew(br) - 3 * (4 / br);
    // an empty workgroup always takes a segment of at least kMinRows rows,
    // so every problem is placed
    B = std::max<long>(B, kFill + (kMinRows + br - 1) / br);
    segs.clear();
    wgseg.assign(1, 0);
    long left_budget = B;
    auto close_wg = [&]() {
        wgseg.push_back((int)segs.size());
        left_budget = B;
    };
    for (int p = 0; p < (int)rows.size(); ++p) {
        int a = 2, left = rows[p];
        while (left > 0) {
            const long fit = br * (left_budget - kFill);
            if (left <= fit) {
                const int L = (left + br - 1) / br * br;
                segs.push_back({p, a, L});
                left_budget -= L / br + kFill;
                left = 0;
                if (left_budget - kFill < kMinRows / br) close_wg();
            } else if (fit >= kMinRows) {
                segs.push_back({p, a, (int)fit});
                a += (int)fit;
                left -= (int)fit;
                close_wg();
            } else {
                close_wg();
            }
        }
    }
    if (wgseg.back() != (int)segs.size()) wgseg.push_back((int)segs.size());
}

static void plan_segments(x2g_ctx *ctx, const std::vector<int> &rows, int wpc, int br, std::vector<SweepSeg> &segs,
                          std::vector<int> &wgseg)
{
    segs.clear();
    wgseg.clear();
    const int forced = env_strip_l();
    if (forced > 0) {
        const int L = (forced + br - 1) / br * br;
        for (int p = 0; p < (int)rows.size(); ++p)
            for (int a = 2; a < 2 + rows[p]; a += L) {
                wgseg.push_back((int)segs.size());
                segs.push_back({p, a, L});
            }
        wgseg.push_back((int)segs.size());
        return;
    }
    long R = 0;
    for (int r : rows) R += r;
    // small batches: down to 8-row segments (a shorter pipeline per
    // workgroup pays more than the halo re-reads cost while CUs are idle)
    const long nwg = std::max<long>(1, std::min<long>((long)wpc * ctx->num_cu, (R + 7) / 8));
    // a budget of all rows plus a fill per problem fits in one workgroup
    const long fill = 32 / br + 4 * sweep_skew(br) - 3 * (4 / br);
    long lo = fill + 8 / br, hi = R / br + (fill + 1) * (long)rows.size() + fill + 8 / br;
    while (lo < hi) {
        const long mid = (lo + hi) / 2;
        segments_for_budget(rows, mid, br, segs, wgseg);
        if ((long)wgseg.size() - 1 <= nwg) hi = mid; else lo = mid + 1;
    }
    segments_for_budget(rows, lo, br, segs, wgseg);
}

// kind: 0 per-pass kernels (in place), 1 k_strip, 2 k_sweep (residue-split lhs/rhs)
template <typename T>
struct SweepPlan {
    std::vector<SweepDesc<T>> even, odd;  // src=A,dst=B and src=B,dst=A
    int kind = 0;
    int skew = 0;
    int total_wg = 0;
    size_t lds = 0;
    // k_sweep: one launch per ring-stride class of the batch's problems (a
    // narrow grid's ring then leaves room for more workgroups per CU than
    // the batch's widest grid would), each with its row segments and every
    // workgroup's first segment (wgseg[w] .. wgseg[w+1]-1; segments index
    // the batch's descriptors), and their staged device copies
    struct Sub {
        int W = 0, br = 4, nth = 0, MS = 0, NT = 0, total_wg = 0;
        size_t lds = 0;
        std::vector<SweepSeg> segs;
        std::vector<int> wgseg;
        const SweepSeg *d_segs = nullptr;
        const int *d_wgseg = nullptr;
    };
    std::vector<Sub> subs;
};

// Does k_sweep apply to a batch whose widest grid has nni points?
// Ring row stride of k_sweep for a batch whose widest grid has nni points:
// the smallest compiled class >= nni + 16 (0: none).  FP64 is compiled for
// the classes up to 352 (wider FP64 rings do not fit the LDS).
template <typename T>
static int sweep_stride(int maxnni)
{
    static const int cls32[] = {128, 192, 256, 384, 528, 768};
    static const int cls64[] = {128, 192, 256, 352};
    const int *c = sizeof(T) == 4 ? cls32 : cls64;
    const int nc = sizeof(T) == 4 ? 6 : 4;
    for (int i = 0; i < nc; ++i)
        if (c[i] >= maxnni + 16) return c[i];
    return 0;
}

// Lanes of a k_sweep launch for a batch whose widest grid has nni points:
// five pass waves with NT colour points per lane and row (64 NT per pass),
// plus spill waves for the MS remaining points of each pass, one per lane.
static void sweep_lanes(int maxnni, int &MS, int &NT, int &nth)
{
    const int M = split_m(maxnni);
    NT = M <= 102 ? 1 : 2;
    MS = std::max(0, M - 64 * NT);
    nth = 320 + 64 * ((5 * MS + 63) / 64);
}

template <typename T>
static bool sweep_fits(x2g_ctx *ctx, int maxnni)
{
    int MS, NT, nth;
    sweep_lanes(maxnni, MS, NT, nth);
    const int W = sweep_stride<T>(maxnni);
    if (W == 0 || nth > 512 || (NT == 2) != (W == 768)) return false;
    // a block's 16-byte chunks must fit the threads' copy slots
    const int P = sweep_row_stride(maxnni, sweep_vec<T>());
    if (4 * P / sweep_vec<T>() > sweep_nsv<T>(4) * nth || P > W) return false;
    return sweep_lds_elems(W, nth, 4) * sizeof(T) <= ctx->max_lds;
}

// Which kernel family a batch of this precision and width uses (engine setting
// and LDS limits); the caller lays out lhs / rhs accordingly.
template <typename T>
static int sweep_kind(x2g_ctx *ctx, int maxnni)
{
    const int e = env_engine();
    if (e == 1) return 0;
    if (e == 0 && sweep_fits<T>(ctx, maxnni)) return 2;
    return choose_skew<T>(ctx, maxnni) > 0 ? 1 : 0;
}

// Launch one sweep of every problem in `descs` (device copy dd, host copy hv).
// k_sweep for the plan's stride class and depth (the instantiations are
// listed once here and in set_lds_attr).
template <typename T, int NT, int W, int BR>
static void launch_ks(x2g_ctx *ctx, const SweepDesc<T> *dd, const typename SweepPlan<T>::Sub &plan, int zero_src)
{
    // the caller has flushed the staged descriptors
    hipLaunchKernelGGL((k_sweep<T, NT, W, BR>), dim3(plan.total_wg), dim3(plan.nth), plan.lds, ctx->stream, dd,
                       plan.d_segs, plan.d_wgseg, plan.MS, zero_src);
}

template <typename T, int BR>
static bool dispatch_ks_br(x2g_ctx *ctx, const SweepDesc<T> *dd, const typename SweepPlan<T>::Sub &plan,
                           int zero_src)
{
    switch (plan.W) {
    case 128: launch_ks<T, 1, 128, BR>(ctx, dd, plan, zero_src); return true;
    case 192: launch_ks<T, 1, 192, BR>(ctx, dd, plan, zero_src); return true;
    case 256: launch_ks<T, 1, 256, BR>(ctx, dd, plan, zero_src); return true;
    default: break;
    }
    if constexpr (sizeof(T) == 8) {
        if (plan.W == 352) { launch_ks<T, 1, 352, BR>(ctx, dd, plan, zero_src); return true; }
    }
    if constexpr (sizeof(T) == 4) {
        if (plan.W == 384) { launch_ks<T, 1, 384, BR>(ctx, dd, plan, zero_src); return true; }
        if (plan.W == 528) { launch_ks<T, 1, 528, BR>(ctx, dd, plan, zero_src); return true; }
        if (plan.W == 768) { launch_ks<T, 2, 768, BR>(ctx, dd, plan, zero_src); return true; }
    }
    return false;
}

template <typename T>
static bool dispatch_ks(x2g_ctx *ctx, const SweepDesc<T> *dd, const typename SweepPlan<T>::Sub &plan, int zero_src)
{
    return plan.br == 2 ? dispatch_ks_br<T, 2>(ctx, dd, plan, zero_src) : dispatch_ks_br<T, 4>(ctx, dd, plan, zero_src);
}

// One sweep of every problem in the plan (device descriptors dd, host copy
// hv); zero_src (k_sweep only): every src array is zero, so none is read.
template <typename T>
static int launch_sweep_once(x2g_ctx *ctx, const SweepDesc<T> *dd, const std::vector<SweepDesc<T>> &hv,
                             const SweepPlan<T> &plan, bool zero_src)
{
    const int skew = plan.skew, total_wg = plan.total_wg;
    const size_t lds = plan.lds;
    const bool in_place = plan.kind == 0;
    const int n = (int)hv.size();
    if (plan.kind == 2) {
        for (const auto &sub : plan.subs) {
            const bool ok = dispatch_ks<T>(ctx, dd, sub, zero_src ? 1 : 0);
            X2G_CHECK(ok, "no sweep kernel for ring stride %d", sub.W);
            X2G_HIP(hipGetLastError());
        }
        return X2G_OK;
    }
    if (!in_place && skew > 0) {
        const int nth = 512;
        if (skew == 2) {
            X2G_LAUNCH(ctx, (k_strip<T, 2>), dim3(total_wg), dim3(nth), lds, dd, n);
        } else {
            X2G_LAUNCH(ctx, (k_strip<T, 1>), dim3(total_wg), dim3(nth), lds, dd, n);
        }
        X2G_HIP(hipGetLastError());
        return X2G_OK;
    }
    long maxpts = 0;
    for (auto &d : hv) maxpts = std::max<long>(maxpts, (long)(d.nni - 2) * (d.nmu - 5));
    for (int pass = 0; pass < 5; ++pass) {
        X2G_LAUNCH(ctx, (k_phase<T>), dim3(grid_x_for(maxpts, 256), n), dim3(256), 0, dd, pass);
        X2G_HIP(hipGetLastError());
    }
    long maxb = 0;
    for (auto &d : hv) maxb = std::max<long>(maxb, (long)(d.nni - 2) + 2L * (d.nmu - 4));
    X2G_LAUNCH(ctx, (k_boundary<T>), dim3(grid_x_for(maxb, 256), n), dim3(256), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

// nsweeps back-to-back sweeps, alternating the even (src = A) and odd
// (src = B) descriptor lists.  With statistics on, one pair of HIP events
// brackets the whole run, so the per-launch average includes the short
// launch-to-launch gaps and no event sits between two sweeps.  a_zero: the
// A arrays are zero (a correction's start), which the first k_sweep launch
// takes without reading them.
template <typename T>
static int launch_sweeps(x2g_ctx *ctx, const SweepDesc<T> *de, const SweepDesc<T> *dodd, const SweepPlan<T> &plan,
                         int nsweeps, bool a_zero = false)
{
    if (nsweeps <= 0) return X2G_OK;
    // algorithmic bytes of one sweep: read u, write u, read lhs and rhs per
    // relaxed point (b and d are 1-D and neighbour reads are reuse)
    double bytes = 0.0;
    for (auto &d : plan.even) bytes += 4.0 * sizeof(T) * (double)(d.nni - 2) * (double)(d.nmu - 5);
    const int prec = sizeof(T) == 8 ? 0 : 1;
    // descriptors go out before the timed region
    int rc = stage_flush(ctx);
    if (rc) return rc;
    hipEvent_t ea, eb;
    if ((rc = stats_begin(ctx, &ea, &eb))) return rc;
    double total = 0.0;
    for (int it = 0; it < nsweeps; ++it) {
        const bool ev = (it % 2) == 0;
        const bool zsrc = it == 0 && a_zero && plan.kind == 2;
        if ((rc = launch_sweep_once<T>(ctx, ev ? de : dodd, ev ? plan.even : plan.odd, plan, zsrc))) return rc;
        total += zsrc ? 0.75 * bytes : bytes;  // a zero source is not read: 3 words per point
    }
    return stats_end(ctx, ea, eb, total, prec, nsweeps);
}

static std::mutex g_attr_mutex;
static int set_lds_attr(x2g_ctx *ctx)
{
    static bool done = false;
    std::lock_guard<std::mutex> lock(g_attr_mutex);
    if (done) return X2G_OK;
    const int mx = (int)ctx->max_lds;
    X2G_HIP(hipFuncSetAttribute((const void *)k_strip<double, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
    X2G_HIP(hipFuncSetAttribute((const void *)k_strip<double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
    X2G_HIP(hipFuncSetAttribute((const void *)k_strip<float, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
    X2G_HIP(hipFuncSetAttribute((const void *)k_strip<float, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
    const void *ks[] = {
        (const void *)k_sweep<double, 1, 128, 4>,
        (const void *)k_sweep<double, 1, 192, 4>,
        (const void *)k_sweep<double, 1, 256, 4>,
        (const void *)k_sweep<double, 1, 352, 4>,
        (const void *)k_sweep<float, 1, 128, 4>,
        (const void *)k_sweep<float, 1, 192, 4>,
        (const void *)k_sweep<float, 1, 256, 4>,
        (const void *)k_sweep<float, 1, 384, 4>,
        (const void *)k_sweep<float, 1, 528, 4>,
        (const void *)k_sweep<float, 2, 768, 4>,
        (const void *)k_sweep<double, 1, 128, 2>,
        (const void *)k_sweep<double, 1, 192, 2>,
        (const void *)k_sweep<double, 1, 256, 2>,
        (const void *)k_sweep<double, 1, 352, 2>,
        (const void *)k_sweep<float, 1, 128, 2>,
        (const void *)k_sweep<float, 1, 192, 2>,
        (const void *)k_sweep<float, 1, 256, 2>,
        (const void *)k_sweep<float, 1, 384, 2>,
        (const void *)k_sweep<float, 1, 528, 2>,
        (const void *)k_sweep<float, 2, 768, 2>};
    for (const void *k : ks) X2G_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
    done = true;
    return X2G_OK;
}

template <typename T>
static void fill_desc(SweepDesc<T> &d, x2g_grid *g, const T *lhs, const T *rhs, const T *bmu,
                      const T *dnu, int isym, double omega)
{
    for (int i = 0; i < 16; ++i) d.c[i] = (T)g->coef[i];
    for (int i = 0; i < 5; ++i) d.ex[i] = (T)g->coef[16 + i];
    d.omega = (T)omega;
    d.omega1 = (T)(1.0 - omega);
    d.lhs = lhs;
    d.rhs = rhs;
    d.bmu = bmu;
    d.dnu = dnu;
    d.nni = g->nni;
    d.nmu = g->nmu;
    d.isym = isym;
    d.cA = colour_mult(g->nni);
    d.wg_begin = 0;
    d.L = 0;
}

// Row stride of the padded arrays that a sweep of this kind works on.
template <typename T>
static int pad_row_stride(int nni, int kind)
{
    return kind == 2 ? sweep_row_stride(nni, sweep_vec<T>()) : nni + 8;
}

// Build the even/odd descriptor lists of a batch; a_bufs/b_bufs are the
// ping-pong padded arrays, with the row stride pad_row_stride<T>(nni, kind)
// of the batch's kind sweep_kind<T>(ctx, widest nni).
template <typename T>
static int plan_sweeps(x2g_ctx *ctx, std::vector<SweepDesc<T>> &base, std::vector<T *> &a_bufs,
                       std::vector<T *> &b_bufs, SweepPlan<T> &plan, const SweepKnobs *kn = nullptr)
{
    int maxnni = 0;
    long rows = 0;
    std::vector<int> prow;
    for (auto &d : base) {
        maxnni = std::max(maxnni, d.nni);
        rows += d.nmu - 5;
        prow.push_back(d.nmu - 5);
    }
    plan.kind = sweep_kind<T>(ctx, maxnni);
    plan.subs.clear();
    int L = 0;
    if (plan.kind == 2) {
        plan.skew = 2;
        L = 0;  // the row segments below
    } else if (plan.kind == 1) {
        plan.skew = choose_skew<T>(ctx, maxnni);
        L = choose_L(ctx, rows);
        plan.lds = strip_lds<T>(maxnni, plan.skew);
    } else {
        plan.skew = 0;
        L = choose_L(ctx, rows);
        plan.lds = 0;
    }
    int wg = 0;
    for (auto &d : base) {
        d.L = L;
        d.wg_begin = wg;
        if (L > 0) wg += (d.nmu - 5 + L - 1) / L;
        d.P = pad_row_stride<T>(d.nni, plan.kind);
        d.pad_bytes = (unsigned)((size_t)d.P * (d.nmu + 8) * sizeof(T));
    }
    plan.total_wg = wg;
    if (plan.kind == 2) {
        // one launch per ring class (all problems in the widest grid's class
        // when the split is off), each planned to fill the chip
        std::map<int, std::vector<int>> cls;  // W -> problem indices
        for (int i = 0; i < (int)base.size(); ++i)
            cls[(kn ? kn->split : g_sweep_split) ? sweep_stride<T>(base[i].nni) : sweep_stride<T>(maxnni)]
                .push_back(i);
        plan.total_wg = 0;
        for (auto &kv : cls) {
            typename SweepPlan<T>::Sub sub;
            int cmax = 0;
            std::vector<int> rows;
            for (int i : kv.second) {
                cmax = std::max(cmax, base[i].nni);
                rows.push_back(prow[i]);
            }
            sweep_lanes(cmax, sub.MS, sub.NT, sub.nth);
            sub.W = kv.first;
            sub.br = (kn ? kn->br : g_sweep_br) == 2 ? 2 : 4;
            sub.lds = sweep_lds_elems(sub.W, sub.nth, sub.br) * sizeof(T);
            int wpc = kn ? kn->wpc : g_sweep_wpc;
            if (wpc <= 0) wpc = std::max<int>(1, (int)(ctx->max_lds / std::max<size_t>(sub.lds, 1)));
            plan_segments(ctx, rows, wpc, sub.br, sub.segs, sub.wgseg);
            for (auto &sg : sub.segs) sg.prob = kv.second[sg.prob];  // class index -> batch index
            sub.total_wg = (int)sub.wgseg.size() - 1;
            plan.total_wg += sub.total_wg;
            plan.lds = std::max(plan.lds, sub.lds);
            plan.subs.push_back(std::move(sub));
        }
    }
    plan.even = base;
    plan.odd = base;
    for (size_t i = 0; i < base.size(); ++i) {
        if (plan.kind != 0) {
            plan.even[i].src = a_bufs[i];
            plan.even[i].dst = b_bufs[i];
            plan.odd[i].src = b_bufs[i];
            plan.odd[i].dst = a_bufs[i];
        } else {
            plan.even[i].src = plan.even[i].dst = a_bufs[i];
            plan.odd[i].src = plan.odd[i].dst = a_bufs[i];
        }
    }
    return X2G_OK;
}

// Stage a plan's even/odd descriptors and, for k_sweep, its segment lists.
template <typename T>
static int stage_plan(x2g_ctx *ctx, SweepPlan<T> &plan, const SweepDesc<T> **de, const SweepDesc<T> **dodd)
{
    size_t off = 0;
    int rc;
    if ((rc = stage_descs(ctx, plan.even, off, de))) return rc;
    if ((rc = stage_descs(ctx, plan.odd, off, dodd))) return rc;
    if (plan.kind == 2) {
        X2G_CHECK(!plan.subs.empty(), "sweep plan without launches");
        for (auto &sub : plan.subs) {
            X2G_CHECK(!sub.segs.empty() && sub.total_wg > 0, "sweep plan without segments");
            if ((rc = stage_descs(ctx, sub.segs, off, &sub.d_segs))) return rc;
            if ((rc = stage_descs(ctx, sub.wgseg, off, &sub.d_wgseg))) return rc;
        }
    }
    return X2G_OK;
}

// ---------------------------------------------------------------------------
// Autotuner of the k_sweep plan.  On the first batch of a shape (precision,
// ring class of the widest grid, number of ring classes, total relaxed rows
// rounded to a power of sqrt 2) every candidate (workgroups per CU 1, 2, 3
// or as many as the LDS allows; one launch or one per ring class) sweeps
// the batch's own buffers a few times under HIP events, and the fastest is
// kept for that shape for the rest of the process.  The candidates' sweeps
// run before the relaxation proper, on work arrays that it rewrites before
// reading (putin, and the zeroed correction arrays), so results do not
// depend on tuning.
struct TuneEntry {
    SweepKnobs best;
    std::vector<std::pair<SweepKnobs, float>> tried;  // candidate, ms per sweep
    long uses = 0;
};
static std::mutex g_tune_mutex;
static std::map<std::string, TuneEntry> g_tune;

template <typename T>
static std::string tune_key(const std::vector<SweepDesc<T>> &base)
{
    int maxnni = 0;
    long rows = 0;
    std::map<int, int> cls;
    for (auto &d : base) {
        maxnni = std::max(maxnni, d.nni);
        rows += d.nmu - 5;
        cls[sweep_stride<T>(d.nni)]++;
    }
    const int bucket = (int)std::lround(2.0 * std::log2((double)std::max<long>(rows, 1)));
    char buf[96];
    snprintf(buf, sizeof(buf), "fp%d W%d classes%d rows~%ld", (int)(8 * sizeof(T)), sweep_stride<T>(maxnni),
             (int)cls.size(), (long)std::lround(std::pow(2.0, bucket / 2.0)));
    return buf;
}

// The knobs for this batch (tuning it if its shape is new).  zsrc: the first
// sweep takes its source as zero (FP32 corrections).
template <typename T>
static int tuned_knobs(x2g_ctx *ctx, std::vector<SweepDesc<T>> &base, std::vector<T *> &A, std::vector<T *> &B,
                       bool zsrc, SweepKnobs &out)
{
    engine_defaults();
    out = SweepKnobs{g_sweep_wpc, g_sweep_split, g_sweep_br};
    if (!g_autotune || g_plan_fixed || env_strip_l() > 0) return X2G_OK;
    int maxnni = 0;
    for (auto &d : base) maxnni = std::max(maxnni, d.nni);
    if (sweep_kind<T>(ctx, maxnni) != 2) return X2G_OK;
    const std::string key = tune_key<T>(base);
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            out = it->second.best;
            it->second.uses++;
            return X2G_OK;
        }
    }
    std::vector<SweepKnobs> cand;
    std::map<int, int> cls;
    for (auto &d : base) cls[sweep_stride<T>(d.nni)]++;
    for (int br : {4, 2})
        for (int split = 0; split <= (cls.size() > 1 ? 1 : 0); ++split)
            for (int wpc : {0, 1, 2, 3}) cand.push_back({wpc, split, br});
    TuneEntry e;
    float best_ms = 1e30f;
    hipEvent_t ea, eb;
    X2G_HIP(hipEventCreate(&ea));
    X2G_HIP(hipEventCreate(&eb));
    std::vector<std::vector<int>> seen;  // block height and workgroups per launch of the plans tried
    int rc = X2G_OK;
    for (const auto &kn : cand) {
        SweepPlan<T> plan;
        std::vector<SweepDesc<T>> b2(base);
        plan_sweeps(ctx, b2, A, B, plan, &kn);
        std::vector<int> sig{kn.br};
        for (auto &sub : plan.subs) sig.push_back(sub.total_wg);
        if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;  // same plan as an earlier one
        seen.push_back(sig);
        const SweepDesc<T> *de, *dodd;
        if ((rc = stage_plan(ctx, plan, &de, &dodd))) break;
        if ((rc = stage_flush(ctx))) break;
        // one untimed sweep, then three timed ones
        if ((rc = launch_sweep_once<T>(ctx, de, plan.even, plan, zsrc))) break;
        if (hipEventRecord(ea, ctx->stream) != hipSuccess) { rc = X2G_EHIP; break; }
        for (int k = 0; k < 3 && !rc; ++k) rc = launch_sweep_once<T>(ctx, (k % 2) ? de : dodd,
                                                                    (k % 2) ? plan.even : plan.odd, plan, false);
        if (rc) break;
        if (hipEventRecord(eb, ctx->stream) != hipSuccess || hipEventSynchronize(eb) != hipSuccess) {
            rc = X2G_EHIP;
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ea, eb);
        ms /= 3.0f;
        e.tried.push_back({kn, ms});
        if (ms < best_ms) {
            best_ms = ms;
            e.best = kn;
        }
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (rc) {
        set_error("sweep autotuner: a candidate launch failed");
        return rc;
    }
    e.uses = 1;
    out = e.best;
    if (getenv("X2G_TUNE_LOG")) {
        fprintf(stderr, "x2g autotune %s:", key.c_str());
        for (auto &t : e.tried)
            fprintf(stderr, " (br %d wpc %d split %d) %.1f us", t.first.br, t.first.wpc, t.first.split, 1e3 * t.second);
        fprintf(stderr, " -> br %d wpc %d split %d\n", e.best.br, e.best.wpc, e.best.split);
    }
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tune.emplace(key, std::move(e));
    return X2G_OK;
}

// Residue-split copies of lhs and rhs for a k_sweep batch.
template <typename TO>
static int launch_pack(x2g_ctx *ctx, const std::vector<PackDesc<TO>> &v)
{
    if (v.empty()) return X2G_OK;
    const PackDesc<TO> *dd;
    int rc = stage_n(ctx, v.data(), v.size(), &dd);
    if (rc) return rc;
    long maxpts = 0;
    for (auto &d : v) maxpts = std::max<long>(maxpts, (long)(d.nni - 2) * (d.nmu - 5));
    X2G_LAUNCH(ctx, (k_pack<TO>), dim3(grid_x_for(maxpts, 256), (int)v.size()), dim3(256), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

static inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static int launch_restride(x2g_ctx *ctx, const std::vector<RestrideDesc> &v)
{
    const RestrideDesc *dd;
    int rc = stage_n(ctx, v.data(), v.size(), &dd);
    if (rc) return rc;
    long maxe = 0;
    for (auto &d : v) maxe = std::max<long>(maxe, (long)d.ds * d.rows);
    X2G_LAUNCH(ctx, k_restride, dim3(grid_x_for(maxe, 256), (int)v.size()), dim3(256), 0, dd);
    X2G_HIP(hipGetLastError());
    return X2G_OK;
}

}  // namespace x2g

using namespace x2g;

// ---------------------------------------------------------------------------
// Ordered (MC)SOR: the reference's sequential point order, exactly.
//
// sor relaxes the points one after another in the mesh ordering; mcsor (one
// thread) in five passes over every fifth position of that ordering
// (sormcsor.F90:20-132, 143-386).  A point reads the new value of a
// neighbour visited before it and the old value of one visited after it, so
// any schedule that updates a point after every earlier-visited neighbour
// and before every later-visited one reproduces the sequential result bit
// for bit.  The host gives every point a level (0 if no earlier-visited
// neighbour shares its pass, else one more than the largest such
// neighbour's level), and a sweep runs the passes in order, each pass its
// levels in order; the points of one level are independent.  For sor in
// col-wise order the levels are the anti-diagonals nu + mu; mcsor in
// col-wise order needs one level per pass, in middle order a few more where
// the ordering folds back.  Backward sor (the altsweeps option) is sor over
// the reversed sequence.
//
// Relaxation sequence of mesh ordering `ordering` (the reference's mesh,
// initCBAllocArrays.F90:2458-2810) as unpadded indices (mu-1)*nni + nu-1.
// Padded rows i = 6..nmu and columns j = 6..nni+3 are walked in one of four
// patterns: rows outer or columns outer, each index either straight,
// reversed, or "mirrored" (from the middle index down to 6, then from the
// one after the middle up to the end).  middlemt mirrors around one row and
// column further and walks its upper rows from the end back to the middle.
static void mesh_sequence(int nni, int nmu, int ordering, std::vector<int> &seq)
{
    seq.clear();
    seq.reserve((size_t)(nni - 2) * (nmu - 5));
    const int rmid = (nmu + 7) / 2, cmid = (nni + 7) / 2;  // padded middle row / column
    const int rlast = nmu, clast = nni + 3;
    // index k (0-based) of a walk of the rows or columns 6..last
    auto straight = [](int k, int) { return 6 + k; };
    auto reversed = [](int k, int last) { return last - k; };
    auto mirrored = [](int mid) {
        return [mid](int k, int) { const int i = 6 + k; return i <= mid ? mid - i + 6 : i; };
    };
    auto emit = [&](int row, int col) { seq.push_back((row - 5) * nni + (col - 5)); };
    const int nr = rlast - 5, nc = clast - 5;
    auto rows_outer = [&](auto rw, auto cw) {
        for (int a = 0; a < nr; ++a)
            for (int b = 0; b < nc; ++b) emit(rw(a, rlast), cw(b, clast));
    };
    auto cols_outer = [&](auto rw, auto cw) {
        for (int b = 0; b < nc; ++b)
            for (int a = 0; a < nr; ++a) emit(rw(a, rlast), cw(b, clast));
    };
    switch (ordering) {
    case X2G_ORDER_COLWISE: rows_outer(straight, straight); break;
    case X2G_ORDER_MIDDLE: rows_outer(mirrored(rmid), mirrored(cmid)); break;
    case X2G_ORDER_MIDDLEMT: {
        const int rm = rmid + 1;
        auto rw = [rm](int k, int last) { const int i = 6 + k; return i <= rm ? rm - i + 6 : last - (i - rm - 1); };
        rows_outer(rw, mirrored(cmid + 1));
        break;
    }
    case X2G_ORDER_ROWWISE: cols_outer(straight, straight); break;
    case X2G_ORDER_RROWWISE: cols_outer(reversed, reversed); break;
    case X2G_ORDER_TEST1:
    case X2G_ORDER_TEST3:
    case X2G_ORDER_MIDDLE2: cols_outer(mirrored(rmid), mirrored(cmid)); break;
    case X2G_ORDER_TEST2: cols_outer(reversed, mirrored(cmid)); break;
    case X2G_ORDER_TEST4: rows_outer(reversed, mirrored(cmid)); break;
    default: break;
    }
}

static void build_levels(int nni, int nmu, int method, int ordering, std::vector<int> &off,
                         std::vector<int> &pts)
{
    std::vector<int> seq;
    mesh_sequence(nni, nmu, ordering, seq);
    if (method == X2G_SOR_BACKWARD) std::reverse(seq.begin(), seq.end());
    const int n = (int)seq.size(), npass = method == X2G_MCSOR ? 5 : 1;
    std::vector<int> pos((size_t)nni * nmu, -1), lvl((size_t)nni * nmu, 0);
    for (int k = 0; k < n; ++k) pos[seq[k]] = k;
    off.assign(1, 0);
    pts.clear();
    std::vector<int> pl, cnt;
    for (int c = 0; c < npass; ++c) {
        pl.clear();
        for (int k = c; k < n; k += npass) pl.push_back(seq[k]);
        int maxl = 0;
        for (int ik : pl) {
            const int mu = ik / nni + 1, nu = ik % nni + 1, me = pos[ik];
            int l = 0;
            auto see = [&](int nu2, int mu2) {
                if (nu2 < 2 || nu2 > nni - 1 || mu2 < 2 || mu2 > nmu - 4) return;   // not relaxed
                const int q = (mu2 - 1) * nni + nu2 - 1, pq = pos[q];
                if (pq < me && (pq % npass) == c) l = std::max(l, lvl[q] + 1);
            };
            for (int k = 1; k <= 4; ++k) {
                see(nu - k, mu);
                see(nu + k, mu);
                see(nu, mu - k);
                see(nu, mu + k);
            }
            lvl[ik] = l;
            maxl = std::max(maxl, l);
        }
        // the pass's points by level, in sequence order within a level
        cnt.assign(maxl + 2, 0);
        for (int ik : pl) ++cnt[lvl[ik] + 1];
        for (int l = 0; l <= maxl; ++l) cnt[l + 1] += cnt[l];
        const size_t base = pts.size();
        pts.resize(base + pl.size());
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int ik : pl) pts[base + fill[lvl[ik]]++] = ik;
        for (int l = 0; l <= maxl; ++l) off.push_back((int)base + cnt[l + 1]);
    }
}

// The pipelined schedule of the same sweep.  Every operation of one sweep --
// the point updates in the order the sequential routine makes them, then the
// boundary extrapolations of sormcsor.F90:94-129 (the mu = 1 row, then the
// nu = 1 and nu = nni columns) -- gets a level one above every earlier
// operation it conflicts with (one reads what the other writes): the
// updates of the four nearest points each way along nu and mu, and for an
// extrapolation the five points it reads and the points that read the value
// it writes.  Sweep s then runs level l at step l + s*lag.  Two conflicting
// operations u, v of consecutive sweeps stay in order when
// lag > |level(u) - level(v)|; their locations lie at most five points apart
// along nu or mu, so lag = 1 + the largest level difference of two written
// locations that close (an upper bound of the conflicting pairs' maximum).
// Within a step no two operations conflict, and every operation reads the
// values the sequential order gives it: the result is the sequential one.
// Codes in ops: an unpadded point index ik >= 0 for a point update, else
// -1 - (kind << 16 | index) for an extrapolation: kind 0 the mu = 1 row at
// nu = index, kind 1 / 2 the nu = 1 / nu = nni column at mu = index.
static void build_pipe(int nni, int nmu, int method, int ordering, std::vector<int> &off, std::vector<int> &ops,
                       int &lag)
{
    std::vector<int> seq;
    mesh_sequence(nni, nmu, ordering, seq);
    if (method == X2G_SOR_BACKWARD) std::reverse(seq.begin(), seq.end());
    const int n = (int)seq.size(), npass = method == X2G_MCSOR ? 5 : 1;
    // level of the operation writing (nu, mu) in this sweep, -1: not written
    std::vector<int> lv((size_t)(nni + 2) * (nmu + 2), -1);
    auto at = [&](int nu, int mu) -> int & { return lv[(size_t)mu * (nni + 2) + nu]; };
    std::vector<std::pair<int, int>> all;  // (level, code)
    all.reserve((size_t)n + nni + 2 * nmu);
    for (int c = 0; c < npass; ++c)
        for (int k = c; k < n; k += npass) {
            const int ik = seq[k], mu = ik / nni + 1, nu = ik % nni + 1;
            int l = 0;
            for (int j = 1; j <= 4; ++j) {
                if (nu - j >= 1) l = std::max(l, at(nu - j, mu) + 1);
                if (nu + j <= nni) l = std::max(l, at(nu + j, mu) + 1);
                if (mu - j >= 1) l = std::max(l, at(nu, mu - j) + 1);
                if (mu + j <= nmu) l = std::max(l, at(nu, mu + j) + 1);
            }
            at(nu, mu) = l;
            all.push_back({l, ik});
        }
    // an extrapolation reads its five points; the points that read its value
    // are among them
    for (int nu = 2; nu <= nni - 1; ++nu) {
        int l = 0;
        for (int m = 2; m <= 6; ++m) l = std::max(l, at(nu, m) + 1);
        at(nu, 1) = l;
        all.push_back({l, -1 - nu});
    }
    for (int side = 1; side <= 2; ++side)
        for (int mu = 1; mu <= nmu - 4; ++mu) {
            const int nb = side == 1 ? 1 : nni, dir = side == 1 ? 1 : -1;
            int l = 0;
            for (int k = 1; k <= 5; ++k) l = std::max(l, at(nb + k * dir, mu) + 1);
            at(nb, mu) = l;
            all.push_back({l, -1 - ((side << 16) | mu)});
        }
    int D = 0, maxl = 0;
    for (int mu = 1; mu <= nmu; ++mu)
        for (int nu = 1; nu <= nni; ++nu) {
            const int a = at(nu, mu);
            if (a < 0) continue;
            maxl = std::max(maxl, a);
            for (int k = 1; k <= 5; ++k) {
                if (nu + k <= nni && at(nu + k, mu) >= 0) D = std::max(D, std::abs(a - at(nu + k, mu)));
                if (mu + k <= nmu && at(nu, mu + k) >= 0) D = std::max(D, std::abs(a - at(nu, mu + k)));
            }
        }
    lag = D + 1;
    // by level, in the order made above within a level
    std::vector<int> cnt(maxl + 2, 0);
    for (auto &o : all) ++cnt[o.first + 1];
    for (int l = 0; l <= maxl; ++l) cnt[l + 1] += cnt[l];
    off.assign(cnt.begin(), cnt.end());
    ops.assign(all.size(), 0);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (auto &o : all) ops[fill[o.first]++] = o.second;
}

// The level schedule of (method, ordering) on grid g, built and uploaded on
// first use.
static int grid_levels(x2g_grid *g, int method, int ordering, const x2g_levels **out)
{
    const int key = method * 16 + ordering;
    auto it = g->levels.find(key);
    if (it == g->levels.end()) {
        std::vector<int> off, pts;
        build_levels(g->nni, g->nmu, method, ordering, off, pts);
        x2g_levels L;
        int rc = x2g_malloc(g->ctx, off.size() * sizeof(int), (void **)&L.off);
        if (!rc) rc = x2g_malloc(g->ctx, pts.size() * sizeof(int), (void **)&L.pts);
        if (!rc) rc = x2g_h2d(g->ctx, L.off, off.data(), off.size() * sizeof(int));
        if (!rc) rc = x2g_h2d(g->ctx, L.pts, pts.data(), pts.size() * sizeof(int));
        if (rc) {
            if (L.off) x2g_free(g->ctx, L.off);
            if (L.pts) x2g_free(g->ctx, L.pts);
            return rc;
        }
        L.nphase = (int)off.size() - 1;
        std::vector<int> poff, pops;
        build_pipe(g->nni, g->nmu, method, ordering, poff, pops, L.lag);
        rc = x2g_malloc(g->ctx, poff.size() * sizeof(int), (void **)&L.poff);
        if (!rc) rc = x2g_malloc(g->ctx, pops.size() * sizeof(int), (void **)&L.pops);
        if (!rc) rc = x2g_h2d(g->ctx, L.poff, poff.data(), poff.size() * sizeof(int));
        if (!rc) rc = x2g_h2d(g->ctx, L.pops, pops.data(), pops.size() * sizeof(int));
        if (rc) {
            for (int *q : {L.off, L.pts, L.poff, L.pops})
                if (q) x2g_free(g->ctx, q);
            return rc;
        }
        L.nlev = (int)poff.size() - 1;
        it = g->levels.emplace(key, L).first;
    }
    *out = &it->second;
    return X2G_OK;
}

struct LevelDesc {
    double *f;                 // padded array (row stride nni+8), relaxed in place
    const double *lhs, *rhs, *bmu, *dnu;
    const int *off, *pts;      // level schedule
    const int *poff, *pops;    // pipelined level schedule
    double c[16], ex[5];
    double omega, omega1;
    int nni, nmu, isym, nphase, nlev, lag;
};

constexpr int kLevelThreads = 512;

// All nsweeps sweeps of one problem per workgroup: the phases of the level
// schedule in order, a barrier after each, then the boundary lines
// (sormcsor.F90:94-129: the mu = 1 row, then the nu = 1 and nu = nni
// columns, whose mu = 1 corners read the new row).  The point update keeps
// the reference's expression and no fused multiply-add is formed, so the
// result is the reference's bit for bit.
__global__ void __launch_bounds__(kLevelThreads) k_levels(const LevelDesc *__restrict__ ld, int nsweeps)
{
#pragma clang fp contract(off)
    const LevelDesc &d = ld[blockIdx.x];
    const int nni = d.nni, nmu = d.nmu, s1 = nni + 8;
    double *f = d.f;
    const double *c = d.c;
    const bool even = d.isym == 1;
    for (int sw = 0; sw < nsweeps; ++sw) {
        for (int ph = 0; ph < d.nphase; ++ph) {
            const int a = d.off[ph], b = d.off[ph + 1];
            for (int k = a + (int)threadIdx.x; k < b; k += blockDim.x) {
                const int ik = d.pts[k];
                const int mu = ik / nni + 1, nu = ik - (mu - 1) * nni + 1;
                const int p = (mu + 3) * s1 + nu + 3;
                const int q1 = s1, q2 = 2 * s1, q3 = 3 * s1, q4 = 4 * s1;
                const double ddmi2 = c[4] * (f[p - q4] + f[p + q4]) + c[5] * (f[p - q3] + f[p + q3]) +
                                     c[6] * (f[p - q2] + f[p + q2]) + c[7] * (f[p - q1] + f[p + q1]);
                const double ddmi1 = c[0] * (f[p - q4] - f[p + q4]) + c[1] * (f[p - q3] - f[p + q3]) +
                                     c[2] * (f[p - q2] - f[p + q2]) + c[3] * (f[p - q1] - f[p + q1]);
                const double ddni2 = c[12] * (f[p - 4] + f[p + 4]) + c[13] * (f[p - 3] + f[p + 3]) +
                                     c[14] * (f[p - 2] + f[p + 2]) + c[15] * (f[p - 1] + f[p + 1]);
                const double ddni1 = c[8] * (f[p - 4] - f[p + 4]) + c[9] * (f[p - 3] - f[p + 3]) +
                                     c[10] * (f[p - 2] - f[p + 2]) + c[11] * (f[p - 1] - f[p + 1]);
                const double cc = ddmi2 + d.bmu[mu - 1] * ddmi1 + ddni2 + d.dnu[nu - 1] * ddni1;
                f[p] = d.omega * (d.rhs[ik] - cc) / d.lhs[ik] + d.omega1 * f[p];
            }
            __syncthreads();
        }
        const double *ex = d.ex;
        for (int nu = 2 + (int)threadIdx.x; nu <= nni - 1; nu += blockDim.x) {
            const int p = 4 * s1 + nu + 3;
            f[p] = even ? ex[0] * f[p + s1] + ex[1] * f[p + 2 * s1] + ex[2] * f[p + 3 * s1] +
                              ex[3] * f[p + 4 * s1] + ex[4] * f[p + 5 * s1]
                        : 0.0;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < 2 * (nmu - 4); t += blockDim.x) {
            const int side = t / (nmu - 4), mu = 1 + t % (nmu - 4);
            const int p = (mu + 3) * s1 + (side == 0 ? 1 : nni) + 3, dir = side == 0 ? 1 : -1;
            f[p] = even ? ex[0] * f[p + dir] + ex[1] * f[p + 2 * dir] + ex[2] * f[p + 3 * dir] +
                              ex[3] * f[p + 4 * dir] + ex[4] * f[p + 5 * dir]
                        : 0.0;
        }
        __syncthreads();
    }
}

constexpr int kPipeThreads = 1024;
constexpr int kPipeChunk = 64;  // sweeps in flight at most: one table entry per lane of a wave

// All nsweeps sweeps of one problem per workgroup on the pipelined schedule:
// step t runs level t - s*lag of every sweep s that has it, then a barrier
// (at most kPipeChunk sweeps form one pipeline; longer runs take several,
// one after another).  The operations of a step are numbered sweep by sweep
// and spread over the threads; the first wave builds the next step's table
// (where each sweep's level starts in pops, and the running count) in LDS
// while the step runs, so a thread finds its operations without walking the
// level offsets.  A thread takes its operations two at a time: both read
// everything they need before either writes (no two operations of a step
// conflict), so their loads are in flight together.  Point updates and
// extrapolations are the expressions of k_levels (no fused multiply-add):
// the result is the sequential routine's bit for bit.
__global__ void __launch_bounds__(kPipeThreads) k_levels_pipe(const LevelDesc *__restrict__ ld, int nsweeps)
{
#pragma clang fp contract(off)
    __shared__ int tb_base[2][kPipeChunk], tb_cum[2][kPipeChunk + 1], tb_n[2];
    const LevelDesc &d = ld[blockIdx.x];
    const int nni = d.nni, s1 = nni + 8, nlev = d.nlev, lag = d.lag;
    double *f = d.f;
    const double *c = d.c, *ex = d.ex;
    const int *poff = d.poff, *pops = d.pops;
    const bool even = d.isym == 1;
    const int lane = threadIdx.x & 63;
    const bool wave0 = threadIdx.x < 64;

    // table of step u of a pipeline of ns sweeps into buffer bf (first wave)
    auto build_table = [&](int u, int ns, int bf) {
        const int shi = min(ns - 1, u / lag);
        const int slo = u < nlev ? 0 : (u - nlev) / lag + 1;
        const int na = shi - slo + 1;
        int base = 0, w = 0;
        if (lane < na) {
            const int l = u - (slo + lane) * lag;
            base = poff[l];
            w = poff[l + 1] - base;
        }
        int x = w;  // inclusive running count over the lanes
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane < na) {
            tb_base[bf][lane] = base;
            tb_cum[bf][lane + 1] = x;
        }
        if (lane == 0) tb_cum[bf][0] = 0;
        if (lane == 63) tb_n[bf] = x;  // lanes past na add nothing
    };
    // value of operation `code` and the padded index it writes
    auto eval = [&](int code, int &pw) -> double {
        if (code >= 0) {
            const int ik = code;
            const int mu = ik / nni + 1, nu = ik - (mu - 1) * nni + 1;
            const int p = (mu + 3) * s1 + nu + 3;
            const int q1 = s1, q2 = 2 * s1, q3 = 3 * s1, q4 = 4 * s1;
            const double ddmi2 = c[4] * (f[p - q4] + f[p + q4]) + c[5] * (f[p - q3] + f[p + q3]) +
                                 c[6] * (f[p - q2] + f[p + q2]) + c[7] * (f[p - q1] + f[p + q1]);
            const double ddmi1 = c[0] * (f[p - q4] - f[p + q4]) + c[1] * (f[p - q3] - f[p + q3]) +
                                 c[2] * (f[p - q2] - f[p + q2]) + c[3] * (f[p - q1] - f[p + q1]);
            const double ddni2 = c[12] * (f[p - 4] + f[p + 4]) + c[13] * (f[p - 3] + f[p + 3]) +
                                 c[14] * (f[p - 2] + f[p + 2]) + c[15] * (f[p - 1] + f[p + 1]);
            const double ddni1 = c[8] * (f[p - 4] - f[p + 4]) + c[9] * (f[p - 3] - f[p + 3]) +
                                 c[10] * (f[p - 2] - f[p + 2]) + c[11] * (f[p - 1] - f[p + 1]);
            const double cc = ddmi2 + d.bmu[mu - 1] * ddmi1 + ddni2 + d.dnu[nu - 1] * ddni1;
            pw = p;
            return d.omega * (d.rhs[ik] - cc) / d.lhs[ik] + d.omega1 * f[p];
        }
        // kind 0: (nu = idx, mu = 1) from mu = 2..6; kind 1 / 2: (nu = 1 /
        // nni, mu = idx) from nu = 2..6 / nni-1..nni-5
        const int b = -1 - code, kind = b >> 16, idx = b & 0xffff;
        const int p = kind == 0 ? 4 * s1 + idx + 3 : (idx + 3) * s1 + (kind == 1 ? 1 : nni) + 3;
        const int st = kind == 0 ? s1 : (kind == 1 ? 1 : -1);
        pw = p;
        return even ? ex[0] * f[p + st] + ex[1] * f[p + 2 * st] + ex[2] * f[p + 3 * st] + ex[3] * f[p + 4 * st] +
                          ex[4] * f[p + 5 * st]
                    : 0.0;
    };

    for (int s0 = 0; s0 < nsweeps; s0 += kPipeChunk) {
        const int ns = min(kPipeChunk, nsweeps - s0);
        const int nstep = nlev + (ns - 1) * lag;
        if (wave0) build_table(0, ns, 0);
        __syncthreads();
        for (int t = 0; t < nstep; ++t) {
            const int bf = t & 1;
            if (wave0 && t + 1 < nstep) build_table(t + 1, ns, bf ^ 1);
            const int total = tb_n[bf];
            const int na = min(ns, t / lag + 1) - (t < nlev ? 0 : (t - nlev) / lag + 1);
            auto code_of = [&](int k) {
                int j = 0;
                for (int i = 1; i < na; ++i) j += k >= tb_cum[bf][i];
                return pops[tb_base[bf][j] + k - tb_cum[bf][j]];
            };
            for (int k = threadIdx.x; k < total; k += 2 * blockDim.x) {
                const int k2 = k + blockDim.x;
                const bool two = k2 < total;
                const int ca = code_of(k), cb = two ? code_of(k2) : ca;
                int pa, pb;
                const double va = eval(ca, pa);
                const double vb = two ? eval(cb, pb) : 0.0;
                f[pa] = va;
                if (two) f[pb] = vb;
            }
            __syncthreads();
        }
    }
}

static void fill_level_desc(LevelDesc &d, x2g_grid *g, const x2g_levels *L, double *f, const double *lhs,
                            const double *rhs, int bkind, int isym, double omega)
{
    d.f = f;
    d.lhs = lhs;
    d.rhs = rhs;
    d.bmu = bkind == 0 ? g->bmu_orb : g->bmu_pot;
    d.dnu = g->dnu;
    d.off = L->off;
    d.pts = L->pts;
    for (int i = 0; i < 16; ++i) d.c[i] = g->coef[i];
    for (int i = 0; i < 5; ++i) d.ex[i] = g->coef[16 + i];
    d.omega = omega;
    d.omega1 = 1.0 - omega;
    d.nni = g->nni;
    d.nmu = g->nmu;
    d.isym = isym;
    d.nphase = L->nphase;
    d.poff = L->poff;
    d.pops = L->pops;
    d.nlev = L->nlev;
    d.lag = L->lag;
}

// nsweeps ordered sweeps of every problem in v (one workgroup each).
static int launch_levels(x2g_ctx *ctx, const std::vector<LevelDesc> &v, int nsweeps)
{
    if (v.empty() || nsweeps <= 0) return X2G_OK;
    const LevelDesc *dd;
    int rc = stage_n(ctx, v.data(), v.size(), &dd);
    if (rc) return rc;
    if ((rc = stage_flush(ctx))) return rc;
    double bytes = 0.0;
    for (auto &d : v) bytes += 32.0 * (double)(d.nni - 2) * (double)(d.nmu - 5) * nsweeps;
    hipEvent_t ea, eb;
    if ((rc = stats_begin(ctx, &ea, &eb))) return rc;
    if (levels_pipe())
        X2G_LAUNCH(ctx, k_levels_pipe, dim3((int)v.size()), dim3(kPipeThreads), 0, dd, nsweeps);
    else
        X2G_LAUNCH(ctx, k_levels, dim3((int)v.size()), dim3(kLevelThreads), 0, dd, nsweeps);
    X2G_HIP(hipGetLastError());
    return stats_end(ctx, ea, eb, bytes, 0, nsweeps);
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" int x2g_set_autotune(int on)
{
    engine_defaults();
    g_autotune = on != 0;
    ++g_plan_gen;
    if (on < 0) {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tune.clear();
        g_autotune = 1;
    }
    return X2G_OK;
}

extern "C" int x2g_set_sweep_plan(int block_rows, int wg_per_cu, int split)
{
    X2G_CHECK(block_rows == 0 || block_rows == 2 || block_rows == 4,
              "x2g_set_sweep_plan: block_rows must be 0, 2 or 4");
    X2G_CHECK(wg_per_cu >= 0 && wg_per_cu <= 16, "x2g_set_sweep_plan: wg_per_cu must be 0..16");
    X2G_CHECK(split == 0 || split == 1, "x2g_set_sweep_plan: split must be 0 or 1");
    engine_defaults();
    g_sweep_br = block_rows ? block_rows : 4;
    g_sweep_wpc = wg_per_cu;
    g_sweep_split = split;
    g_plan_fixed = block_rows || wg_per_cu || split;
    ++g_plan_gen;
    return X2G_OK;
}

extern "C" int x2g_tune_report(char *buf, int len)
{
    X2G_CHECK(buf && len > 0, "x2g_tune_report: null buffer");
    std::string out;
    {
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        for (auto &kv : g_tune) {
            char line[512];
            int n = snprintf(line, sizeof(line), "%s: br %d wpc %d split %d (batches %ld);", kv.first.c_str(),
                             kv.second.best.br, kv.second.best.wpc, kv.second.best.split, kv.second.uses);
            for (auto &t : kv.second.tried)
                if (n < (int)sizeof(line) - 40)
                    n += snprintf(line + n, sizeof(line) - n, " %d/%d/%d %.1fus", t.first.br, t.first.wpc,
                                  t.first.split, 1e3 * t.second);
            out += line;
            out += "\n";
        }
    }
    snprintf(buf, (size_t)len, "%s", out.c_str());
    return (int)out.size();
}

extern "C" int x2g_set_levels_pipeline(int on)
{
    g_levels_pipe = on != 0;
    ++g_plan_gen;
    return X2G_OK;
}

extern "C" int x2g_set_sweep_engine(int engine, int strip_rows)
{
    X2G_CHECK(engine >= 0 && engine <= 2, "x2g_set_sweep_engine: engine must be 0, 1 or 2");
    engine_defaults();
    g_engine = engine;
    g_strip_rows = strip_rows;
    ++g_plan_gen;
    return X2G_OK;
}

extern "C" int x2g_putin(x2g_grid *g, int isym, const double *fun_d, double *work_d)
{
    X2G_CHECK(g && fun_d && work_d, "x2g_putin: null argument");
    X2G_CHECK(isym == 1 || isym == -1, "x2g_putin: isym must be +1 or -1");
    return launch_putin(g, isym, fun_d, work_d, nullptr);
}

extern "C" int x2g_putout(x2g_grid *g, double *fun_d, const double *work_d)
{
    X2G_CHECK(g && fun_d && work_d, "x2g_putout: null argument");
    return launch_putout(g, fun_d, work_d);
}

static int check_grid_sweepable(x2g_grid *g)
{
    X2G_CHECK(g->nni >= 7 && g->nmu >= 11, "grid %dx%d too small for the 17-point stencil", g->nni, g->nmu);
    // the sweep kernels address each array with 32-bit byte offsets below 1 GiB
    X2G_CHECK((size_t)(g->nni + 16) * (g->nmu + 8) * 2 * sizeof(double) < ((size_t)1 << 30),
              "grid %dx%d too large for the sweep kernels", g->nni, g->nmu);
    return X2G_OK;
}

// Static device scratch of x2g_mcsor_noalloc: the workspace of one mcsor
// call (two padded sweep buffers and the colour-ordered lhs/rhs pairs, or a
// ping-pong partner of the caller's array) for callers that cannot allocate.
// Each GPU holds its own copy for the life of the process; host code uses
// only its per-device address from hipGetSymbolAddress (x2g::scratch_init).
constexpr size_t kScratchBytes = (size_t)24 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace x2g {
// The device address of `scratch` on the context's device (which is
// current): called once per context by x2g_ctx_create.
int scratch_init(x2g_ctx *ctx)
{
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error("hipGetSymbolAddress(scratch) on device %d: %s", ctx->device, hipGetErrorString(e));
        return X2G_EHIP;
    }
    ctx->scratch = ptr;
    ctx->scratch_bytes = kScratchBytes;
    return X2G_OK;
}
}  // namespace x2g

// Workspace bytes of one mcsor call of this sweep kind on grid g.
static size_t mcsor_ws_bytes(const x2g_grid *g, int kind)
{
    const int P = pad_row_stride<double>(g->nni, kind);
    const size_t ubytes = g->npad() * sizeof(double);
    const size_t pbytes = (size_t)P * (g->nmu + 8) * sizeof(double);
    const size_t lrbytes = 2 * split_size(g->nni, g->nmu) * sizeof(double);
    return kind == 2 ? 2 * align256(pbytes) + align256(lrbytes) : align256(ubytes);
}

// mcsor of one padded array; the workspace is w0 (at least mcsor_ws_bytes).
// Every workspace byte a kernel reads is written earlier in the same call
// (the restride fills both sweep buffers, k_pack the pairs; absent points of
// a short colour block are computed into an LDS sink), so nothing in it has
// to start at zero.
static int mcsor_run(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d, const double *rhs_d,
                     double omega, int nsweeps, int kind, char *w0)
{
    x2g_ctx *ctx = g->ctx;
    int rc;
    const int P = pad_row_stride<double>(g->nni, kind);
    const size_t ubytes = g->npad() * sizeof(double);                 // the caller's array (stride nni+8)
    const size_t pbytes = (size_t)P * (g->nmu + 8) * sizeof(double);  // a sweep buffer (stride P)
    const size_t lrbytes = 2 * split_size(g->nni, g->nmu) * sizeof(double);
    // k_sweep: two buffers of stride P and the (omega/lhs, rhs) pairs; the
    // other kinds: a ping-pong partner of the caller's array
    const bool restride = kind == 2;
    double *A = work_d, *B = (double *)w0;
    if (restride) {
        A = (double *)w0;
        B = (double *)(w0 + align256(pbytes));
        if ((rc = launch_restride(ctx, {{work_d, A, B, g->nni + 8, g->nmu + 8, g->nni + 8, P}}))) return rc;
    } else {
        X2G_HIP(hipMemcpyAsync(B, work_d, ubytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    std::vector<SweepDesc<double>> base(1);
    fill_desc<double>(base[0], g, lhs_d, rhs_d, bkind == 0 ? g->bmu_orb : g->bmu_pot, g->dnu, isym, omega);
    std::vector<double *> VA{A}, VB{B};
    SweepPlan<double> plan;
    plan_sweeps(ctx, base, VA, VB, plan);
    if (plan.kind == 2) {
        double *lr = (double *)(w0 + 2 * align256(pbytes));
        std::vector<PackDesc<double>> pk{{lhs_d, lr, omega, g->nni, g->nmu, 0}, {rhs_d, lr, 0.0, g->nni, g->nmu, 1}};
        if ((rc = launch_pack(ctx, pk))) return rc;
        for (auto *v : {&plan.even, &plan.odd}) {
            (*v)[0].lr = lr;
            (*v)[0].lr_bytes = (unsigned)lrbytes;
        }
    }
    const SweepDesc<double> *de, *dodd;
    if ((rc = stage_plan(ctx, plan, &de, &dodd))) return rc;
    const bool in_place = plan.kind == 0;
    if ((rc = launch_sweeps<double>(ctx, de, dodd, plan, nsweeps))) return rc;
    double *fin = (!in_place && (nsweeps % 2) == 1) ? B : A;
    if (restride) {
        if ((rc = launch_restride(ctx, {{fin, work_d, nullptr, g->nni + 8, g->nmu + 8, P, g->nni + 8}}))) return rc;
    } else if (fin != work_d) {
        X2G_HIP(hipMemcpyAsync(work_d, fin, ubytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return X2G_OK;
}

static int mcsor_check(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d, const double *rhs_d,
                       double omega, int nsweeps, const char *who)
{
    X2G_CHECK(g && work_d && lhs_d && rhs_d, "%s: null argument", who);
    X2G_CHECK(isym == 1 || isym == -1, "%s: isym must be +1 or -1", who);
    X2G_CHECK(bkind == 0 || bkind == 1, "%s: bkind must be 0 or 1", who);
    X2G_CHECK(nsweeps >= 0, "%s: nsweeps < 0", who);
    X2G_CHECK(omega > 0.0 && omega < 2.0, "%s: omega must lie in (0, 2)", who);
    return check_grid_sweepable(g);
}

extern "C" int x2g_mcsor(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d,
                         const double *rhs_d, double omega, int nsweeps)
{
    int rc = mcsor_check(g, isym, bkind, work_d, lhs_d, rhs_d, omega, nsweeps, "x2g_mcsor");
    if (rc || nsweeps == 0) return rc;
    x2g_ctx *ctx = g->ctx;
    if ((rc = set_lds_attr(ctx))) return rc;
    const int kind = sweep_kind<double>(ctx, g->nni);
    if ((rc = ws_reserve(ctx, 0, mcsor_ws_bytes(g, kind)))) return rc;
    return mcsor_run(g, isym, bkind, work_d, lhs_d, rhs_d, omega, nsweeps, kind, (char *)ctx->ws[0].ptr);
}

extern "C" int x2g_mcsor_noalloc(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d,
                                 const double *rhs_d, double omega, int nsweeps)
{
    int rc = mcsor_check(g, isym, bkind, work_d, lhs_d, rhs_d, omega, nsweeps, "x2g_mcsor_noalloc");
    if (rc || nsweeps == 0) return rc;
    x2g_ctx *ctx = g->ctx;
    X2G_CHECK(ctx->scratch, "x2g_mcsor_noalloc: no scratch address for device %d", ctx->device);
    if ((rc = set_lds_attr(ctx))) return rc;
    const int kind = sweep_kind<double>(ctx, g->nni);
    const size_t need = mcsor_ws_bytes(g, kind);
    X2G_CHECK(need <= ctx->scratch_bytes,
              "x2g_mcsor_noalloc: grid %dx%d needs %zu bytes of scratch, more than its %zu (use x2g_mcsor)", g->nni,
              g->nmu, need, ctx->scratch_bytes);
    // the sweep timing's events would be created on demand: not in this call
    const bool stats = ctx->stats_on;
    ctx->stats_on = false;
    rc = mcsor_run(g, isym, bkind, work_d, lhs_d, rhs_d, omega, nsweeps, kind, (char *)ctx->scratch);
    ctx->stats_on = stats;
    return rc;
}

extern "C" int x2g_relax_ordered(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d,
                                 const double *rhs_d, double omega, int nsweeps, int method, int ordering)
{
    int rc = mcsor_check(g, isym, bkind, work_d, lhs_d, rhs_d, omega, nsweeps, "x2g_relax_ordered");
    if (rc) return rc;
    X2G_CHECK(method == X2G_MCSOR || method == X2G_SOR || method == X2G_SOR_BACKWARD,
              "x2g_relax_ordered: method must be X2G_MCSOR, X2G_SOR or X2G_SOR_BACKWARD");
    X2G_CHECK(ordering >= X2G_ORDER_COLWISE && ordering <= X2G_ORDER_MIDDLE2,
              "x2g_relax_ordered: ordering must be one of X2G_ORDER_*");
    if (nsweeps == 0) return X2G_OK;
    const x2g_levels *L;
    if ((rc = grid_levels(g, method, ordering, &L))) return rc;
    std::vector<LevelDesc> v(1);
    fill_level_desc(v[0], g, L, work_d, lhs_d, rhs_d, bkind, isym, omega);
    return launch_levels(g->ctx, v, nsweeps);
}

extern "C" int x2g_mcsor_host(x2g_grid *g, int isym, int bkind, double *work, const double *lhs,
                              const double *rhs, double omega, int nsweeps)
{
    X2G_CHECK(g && work && lhs && rhs, "x2g_mcsor_host: null argument");
    x2g_ctx *ctx = g->ctx;
    const size_t pb = g->npad() * sizeof(double), nb = g->n() * sizeof(double);
    double *w_d = nullptr, *l_d = nullptr, *r_d = nullptr;
    int rc = X2G_OK;
    if (x2g_malloc(ctx, pb, (void **)&w_d) || x2g_malloc(ctx, nb, (void **)&l_d) ||
        x2g_malloc(ctx, nb, (void **)&r_d)) {
        rc = X2G_ENOMEM;
    } else if (!(rc = x2g_h2d(ctx, w_d, work, pb)) && !(rc = x2g_h2d(ctx, l_d, lhs, nb)) &&
               !(rc = x2g_h2d(ctx, r_d, rhs, nb)) &&
               !(rc = x2g_mcsor(g, isym, bkind, w_d, l_d, r_d, omega, nsweeps))) {
        rc = x2g_d2h(ctx, work, w_d, pb);
    }
    if (w_d) x2g_free(ctx, w_d);
    if (l_d) x2g_free(ctx, l_d);
    if (r_d) x2g_free(ctx, r_d);
    return rc;
}

// An item relaxed in the reference's sequential order (any scheme other than
// the col-wise MCSOR colour sweep).
static inline bool ordered(const x2g_relax_item &it)
{
    return it.method != X2G_MCSOR || it.ordering != X2G_ORDER_COLWISE;
}

static int relax_batch_eager(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_relax_item *items, int n,
                             int maxsor1, int maxsor2)
{
    X2G_CHECK(ctx && grids && (items || n == 0) && n >= 0, "x2g_relax_batch: null argument");
    X2G_CHECK(maxsor1 >= 0 && maxsor2 >= 0, "x2g_relax_batch: negative sweep counts");
    if (n == 0 || maxsor1 == 0) return X2G_OK;
    int rc = set_lds_attr(ctx);
    if (rc) return rc;
    // validate, pick the sweep kind of each precision's batch (it fixes the
    // row stride of the padded arrays), and size the workspace: per FP64
    // item two padded arrays and the (omega/lhs, rhs) pairs; per FP32 item
    // two padded correction arrays and the FP32 lhs / residual.
    int maxnni64 = 0, maxnni32 = 0;
    for (int i = 0; i < n; ++i) {
        const x2g_relax_item &it = items[i];
        X2G_CHECK(it.grid_index >= 0 && grids[it.grid_index], "x2g_relax_batch: bad grid index %d", it.grid_index);
        x2g_grid *g = grids[it.grid_index];
        X2G_CHECK(g->ctx == ctx, "x2g_relax_batch: grid of another context");
        X2G_CHECK(it.fun_d && it.lhs_d && it.rhs_d, "x2g_relax_batch: item %d has a null array", i);
        X2G_CHECK(it.isym == 1 || it.isym == -1, "x2g_relax_batch: item %d isym", i);
        X2G_CHECK(it.bkind == 0 || it.bkind == 1, "x2g_relax_batch: item %d bkind", i);
        X2G_CHECK(it.prec == 0 || it.prec == 1, "x2g_relax_batch: item %d prec", i);
        X2G_CHECK(it.omega > 0.0 && it.omega < 2.0, "x2g_relax_batch: item %d omega must lie in (0, 2)", i);
        X2G_CHECK(it.method == X2G_MCSOR || it.method == X2G_SOR || it.method == X2G_SOR_BACKWARD,
                  "x2g_relax_batch: item %d method", i);
        X2G_CHECK(it.ordering >= X2G_ORDER_COLWISE && it.ordering <= X2G_ORDER_MIDDLE2,
                  "x2g_relax_batch: item %d ordering", i);
        X2G_CHECK(it.prec == 0 || (it.method == X2G_MCSOR && it.ordering == X2G_ORDER_COLWISE),
                  "x2g_relax_batch: item %d: FP32 correction sweeps are col-wise MCSOR only", i);
        if ((rc = check_grid_sweepable(g))) return rc;
        if (!ordered(it)) (it.prec ? maxnni32 : maxnni64) = std::max(it.prec ? maxnni32 : maxnni64, g->nni);
    }
    const int kind64 = maxnni64 ? sweep_kind<double>(ctx, maxnni64) : 0;
    const int kind32 = maxnni32 ? sweep_kind<float>(ctx, maxnni32) : 0;
    auto pstride = [&](int i) {
        x2g_grid *g = grids[items[i].grid_index];
        return items[i].prec ? pad_row_stride<float>(g->nni, kind32) : pad_row_stride<double>(g->nni, kind64);
    };
    auto padbytes = [&](int i) {  // one padded array of item i
        x2g_grid *g = grids[items[i].grid_index];
        if (ordered(items[i])) return g->npad() * sizeof(double);  // relaxed in place, stride nni+8
        return (size_t)pstride(i) * (g->nmu + 8) * (items[i].prec ? sizeof(float) : sizeof(double));
    };
    auto lrbytes = [&](int i) {  // (omega/lhs, rhs) pairs of item i
        x2g_grid *g = grids[items[i].grid_index];
        return (size_t)2 * split_size(g->nni, g->nmu) * (items[i].prec ? sizeof(float) : sizeof(double));
    };
    size_t off64 = 0, off32 = 0;
    std::vector<size_t> o64(n), o32(n);
    for (int i = 0; i < n; ++i) {
        x2g_grid *g = grids[items[i].grid_index];
        o64[i] = off64;
        o32[i] = off32;
        if (ordered(items[i]))
            off64 += align256(padbytes(i));
        else if (items[i].prec == 0)
            off64 += 2 * align256(padbytes(i)) + align256(lrbytes(i));
        else
            off32 += 2 * align256(padbytes(i)) +
                     2 * align256(std::max<size_t>(g->n(), split_size(g->nni, g->nmu)) * sizeof(float));
    }
    if (off64 && (rc = ws_reserve(ctx, 0, off64))) return rc;
    if (off32 && (rc = ws_reserve(ctx, 3, off32))) return rc;
    char *w64 = (char *)ctx->ws[0].ptr;
    char *w32 = (char *)ctx->ws[3].ptr;

    std::vector<int> i64, i32, iord;  // item indices by precision; ordered (MC)SOR items
    for (int i = 0; i < n; ++i) (ordered(items[i]) ? iord : items[i].prec ? i32 : i64).push_back(i);

    auto bufA = [&](int i) { return (double *)(w64 + o64[i]); };
    auto bufB = [&](int i) { return (double *)(w64 + o64[i] + align256(padbytes(i))); };
    auto lrp = [&](int i) { return (double *)(w64 + o64[i] + 2 * align256(padbytes(i))); };  // FP64 pairs
    // FP32 item: k = 0 eA, 1 eB, 2 lhs_f (k_sweep: the (omega/lhs, residual)
    // pairs, which extend over the region of k = 3), 3 res_f
    auto f32p = [&](int i, int k) {
        x2g_grid *g = grids[items[i].grid_index];
        const size_t pp = align256(padbytes(i));
        const size_t nn = align256(std::max<size_t>(g->n(), split_size(g->nni, g->nmu)) * sizeof(float));
        size_t o = o32[i] + (k <= 1 ? k * pp : 2 * pp + (k - 2) * nn);
        return (float *)(w32 + o);
    };

    size_t doff = 0;
    // descriptors that do not change across macro-iterations; FP32 items work
    // on the unpadded function directly (residual with on-the-fly mirrors,
    // in-place correction), so only FP64 items are embedded
    std::vector<PadDesc> pin, pout;
    std::vector<TailDesc> tails;
    for (int i = 0; i < n; ++i) {
        const x2g_relax_item &it = items[i];
        x2g_grid *g = grids[it.grid_index];
        if (ordered(it))
            pin.push_back({it.fun_d, bufA(i), nullptr, g->nni, g->nmu, it.isym, g->nni + 8});
        else if (it.prec == 0)
            pin.push_back({it.fun_d, bufA(i), bufB(i), g->nni, g->nmu, it.isym, pstride(i)});
        if (it.tail_fock1_d)
            tails.push_back({it.fun_d, it.tail_fock1_d, g->suppl + 5 * g->n(), g->vxisq, g->vetasq, g->nni, g->nmu});
    }
    const PadDesc *d_pin = nullptr;
    const TailDesc *d_tail = nullptr;
    if (!pin.empty() && (rc = stage_descs(ctx, pin, doff, &d_pin))) return rc;
    if (!tails.empty() && (rc = stage_descs(ctx, tails, doff, &d_tail))) return rc;

    // FP64 items that give their rhs as a product: form it in rhs_d first
    {
        std::vector<ProdDesc> pv;
        long maxn = 0;
        std::vector<int> prod_items(i64);
        prod_items.insert(prod_items.end(), iord.begin(), iord.end());
        for (int i : prod_items) {
            const x2g_relax_item &it = items[i];
            if (!it.rhs_f1_d) continue;
            const long np = grids[it.grid_index]->n();
            pv.push_back({const_cast<double *>(it.rhs_d), {it.rhs_f1_d, it.rhs_f2_d, it.rhs_f3_d}, np});
            maxn = std::max(maxn, np);
        }
        if (!pv.empty()) {
            const ProdDesc *dp;
            if ((rc = stage_descs(ctx, pv, doff, &dp))) return rc;
            X2G_LAUNCH(ctx, k_rhs_product, dim3(grid_x_for(maxn, 256), (int)pv.size()), dim3(256), 0, dp);
            X2G_HIP(hipGetLastError());
        }
    }

    // FP64 items: sweep plan
    SweepPlan<double> p64;
    const SweepDesc<double> *d64e = nullptr, *d64o = nullptr;
    std::vector<PadDesc> pout64;
    if (!i64.empty()) {
        std::vector<SweepDesc<double>> base(i64.size());
        std::vector<double *> A, B;
        for (size_t k = 0; k < i64.size(); ++k) {
            const x2g_relax_item &it = items[i64[k]];
            x2g_grid *g = grids[it.grid_index];
            fill_desc<double>(base[k], g, it.lhs_d, it.rhs_d, it.bkind == 0 ? g->bmu_orb : g->bmu_pot, g->dnu,
                              it.isym, it.omega);
            A.push_back(bufA(i64[k]));
            B.push_back(bufB(i64[k]));
            base[k].lr = lrp(i64[k]);
            base[k].lr_bytes = (unsigned)lrbytes(i64[k]);
        }
        if (kind64 == 2) {
            // lhs / rhs do not change during the call: one residue-split copy
            std::vector<PackDesc<double>> pk;
            for (size_t k = 0; k < i64.size(); ++k) {
                const x2g_relax_item &it = items[i64[k]];
                x2g_grid *g = grids[it.grid_index];
                pk.push_back({it.lhs_d, lrp(i64[k]), it.omega, g->nni, g->nmu, 0});
                pk.push_back({it.rhs_d, lrp(i64[k]), 0.0, g->nni, g->nmu, 1});
            }
            if ((rc = launch_pack(ctx, pk))) return rc;
        }
        SweepKnobs kn64;
        if ((rc = tuned_knobs<double>(ctx, base, A, B, false, kn64))) return rc;
        plan_sweeps(ctx, base, A, B, p64, &kn64);
        if ((rc = stage_plan(ctx, p64, &d64e, &d64o))) return rc;
        const bool final_in_b = p64.kind != 0 && (maxsor2 % 2) == 1;
        for (size_t k = 0; k < i64.size(); ++k) {
            const x2g_relax_item &it = items[i64[k]];
            x2g_grid *g = grids[it.grid_index];
            pout64.push_back({it.fun_d, final_in_b ? bufB(i64[k]) : bufA(i64[k]), nullptr, g->nni, g->nmu, 1,
                              pstride(i64[k])});
        }
    }
    const PadDesc *d_pout64 = nullptr;
    if (!pout64.empty() && (rc = stage_descs(ctx, pout64, doff, &d_pout64))) return rc;

    // ordered (MC)SOR items: level schedules, putout descriptors
    std::vector<LevelDesc> lev;
    std::vector<PadDesc> pout_ord;
    for (int i : iord) {
        const x2g_relax_item &it = items[i];
        x2g_grid *g = grids[it.grid_index];
        const x2g_levels *L;
        if ((rc = grid_levels(g, it.method, it.ordering, &L))) return rc;
        LevelDesc d{};
        fill_level_desc(d, g, L, bufA(i), it.lhs_d, it.rhs_d, it.bkind, it.isym, it.omega);
        lev.push_back(d);
        pout_ord.push_back({it.fun_d, bufA(i), nullptr, g->nni, g->nmu, 1, g->nni + 8});
    }
    const PadDesc *d_pout_ord = nullptr;
    if (!pout_ord.empty() && (rc = stage_descs(ctx, pout_ord, doff, &d_pout_ord))) return rc;

    // FP32 items: conversion, residual, sweeps, correction
    SweepPlan<float> p32;
    const SweepDesc<float> *d32e = nullptr, *d32o = nullptr;
    const ResDesc *d_res = nullptr;
    const CorrDesc *d_corr = nullptr;
    const CvtDesc *d_cvt = nullptr;
    const FillFDesc *d_zero = nullptr;
    int nzero = 0;
    long maxzero = 0;  // longest zero fill
    bool res_prod = false;  // some FP32 item's rhs is a product of factors
    const BndDesc *d_bnd = nullptr;
    std::vector<BndDesc> bnd;
    if (!i32.empty()) {
        std::vector<SweepDesc<float>> base(i32.size());
        std::vector<float *> A, B;
        std::vector<ResDesc> res;
        std::vector<CorrDesc> corr;
        std::vector<CvtDesc> cvt;
        std::vector<FillFDesc> zero;
        for (size_t k = 0; k < i32.size(); ++k) {
            const int i = i32[k];
            const x2g_relax_item &it = items[i];
            x2g_grid *g = grids[it.grid_index];
            fill_desc<float>(base[k], g, f32p(i, 2), f32p(i, 3), it.bkind == 0 ? g->bmu_orb_f : g->bmu_pot_f,
                             g->dnu_f, it.isym, it.omega);
            A.push_back(f32p(i, 0));
            B.push_back(f32p(i, 1));
            ResDesc r{};
            r.fun = it.fun_d;
            r.isym = it.isym;
            r.lhs = it.lhs_d;
            r.rhs = it.rhs_d;
            r.rf[0] = it.rhs_f1_d;
            r.rf[1] = it.rhs_f1_d ? it.rhs_f2_d : nullptr;
            r.rf[2] = it.rhs_f1_d ? it.rhs_f3_d : nullptr;
            r.bmu = it.bkind == 0 ? g->bmu_orb : g->bmu_pot;
            r.dnu = g->dnu;
            r.res_f = kind32 == 2 ? f32p(i, 2) : f32p(i, 3);
            r.res_split = kind32 == 2;  // k_sweep: (omega/lhs, r) pairs, formed here each macro-iteration
            r.omega = it.omega;
            r.part = nullptr;
            for (int q = 0; q < 16; ++q) r.c[q] = g->coef[q];
            r.nni = g->nni;
            r.nmu = g->nmu;
            res.push_back(r);
            corr.push_back({it.fun_d, f32p(i, 0), g->nni, g->nmu, pstride(i)});
            cvt.push_back({it.lhs_d, f32p(i, 2), (long)g->n()});
            if (kind32 == 2) {
                // the correction starts from zero: the first sweep takes A as
                // zero without reading it, and k_sweep writes every row of
                // its destination except the mu <= 1 and mu >= nmu-3 rows,
                // which stay zero in A and B
                const long P = pstride(i);
                for (int k = 0; k < 2; ++k) {
                    zero.push_back({f32p(i, k), 5 * P});
                    zero.push_back({f32p(i, k) + (long)g->nmu * P, 8 * P});
                }
            } else {
                zero.push_back({f32p(i, 0), (long)(2 * align256(padbytes(i)) / sizeof(float))});
            }
            BndDesc bd{};
            bd.fun = it.fun_d;
            for (int q = 0; q < 5; ++q) bd.ex[q] = g->coef[16 + q];
            bd.nni = g->nni;
            bd.nmu = g->nmu;
            bd.isym = it.isym;
            bnd.push_back(bd);
        }
        if (kind32 == 2)
            for (size_t k = 0; k < i32.size(); ++k) {
                base[k].lr = f32p(i32[k], 2);
                base[k].lr_bytes = (unsigned)lrbytes(i32[k]);
            }
        SweepKnobs kn32;
        if ((rc = tuned_knobs<float>(ctx, base, A, B, true, kn32))) return rc;
        plan_sweeps(ctx, base, A, B, p32, &kn32);
        // the strip engines ping-pong: an odd sweep count leaves the correction in B
        if (p32.kind != 0 && (maxsor2 % 2) == 1)
            for (size_t k = 0; k < i32.size(); ++k) corr[k].e = f32p(i32[k], 1);
        const bool split32 = p32.kind == 2;
        if ((rc = stage_plan(ctx, p32, &d32e, &d32o))) return rc;
        if ((rc = stage_descs(ctx, res, doff, &d_res))) return rc;
        if ((rc = stage_descs(ctx, corr, doff, &d_corr))) return rc;
        if ((rc = stage_descs(ctx, cvt, doff, &d_cvt))) return rc;
        if ((rc = stage_descs(ctx, zero, doff, &d_zero))) return rc;
        nzero = (int)zero.size();
        for (auto &z : zero) maxzero = std::max(maxzero, z.n);
        for (auto &r : res) res_prod = res_prod || r.rf[0];
        if ((rc = stage_descs(ctx, bnd, doff, &d_bnd))) return rc;
        if (!split32) {
            long maxn = 0;
            for (int i : i32) maxn = std::max<long>(maxn, (long)grids[items[i].grid_index]->n());
            X2G_LAUNCH(ctx, k_to_f32, dim3(grid_x_for(maxn, 256), (int)i32.size()), dim3(256), 0,
                               d_cvt);
            X2G_HIP(hipGetLastError());
        }
    }

    long maxpad = 0, maxn = 0, maxnni = 0, maxnmu = 0;
    for (int i = 0; i < n; ++i) {
        x2g_grid *g = grids[items[i].grid_index];
        maxpad = std::max<long>(maxpad, (long)pstride(i) * (g->nmu + 8));
        maxn = std::max<long>(maxn, (long)g->n());
        maxnni = std::max<long>(maxnni, g->nni);
        maxnmu = std::max<long>(maxnmu, g->nmu);
    }
    for (int itr1 = 0; itr1 < maxsor1; ++itr1) {
        // FP64 items: putin into both ping-pong buffers
        if (d_pin) {
            X2G_LAUNCH(ctx, k_putin, dim3(grid_x_for(maxpad, 256), (int)pin.size()), dim3(256), 0,
                               d_pin);
            X2G_HIP(hipGetLastError());
        }
        if (!i64.empty()) {
            if ((rc = launch_sweeps<double>(ctx, d64e, d64o, p64, maxsor2))) return rc;
            X2G_LAUNCH(ctx, k_putout, dim3(grid_x_for(maxn, 256), (int)i64.size()), dim3(256), 0,
                               d_pout64);
            X2G_HIP(hipGetLastError());
        }
        if (!iord.empty() && maxsor2 > 0) {
            if ((rc = launch_levels(ctx, lev, maxsor2))) return rc;
            X2G_LAUNCH(ctx, k_putout, dim3(grid_x_for(maxn, 256), (int)iord.size()), dim3(256), 0,
                               d_pout_ord);
            X2G_HIP(hipGetLastError());
        }
        // (no sweeps: u stays as it is, as in the FP64 path; the correction
        // arrays are not even cleared then)
        if (!i32.empty() && maxsor2 > 0) {
            const int n32 = (int)i32.size();
            if ((rc = launch_residual(ctx, d_res, n32, maxnni, maxnmu, maxn, res_prod))) return rc;
            X2G_LAUNCH(ctx, k_zero_f32, dim3(std::min(grid_x_for(maxzero, 256), 128), nzero), dim3(256), 0, d_zero);
            X2G_HIP(hipGetLastError());
            if ((rc = launch_sweeps<float>(ctx, d32e, d32o, p32, maxsor2, /*a_zero=*/true))) return rc;
            X2G_LAUNCH(ctx, k_correct, dim3((int)std::min<long>(maxnmu, 1024), n32), dim3(256), 0, d_corr);
            X2G_HIP(hipGetLastError());
            X2G_LAUNCH(ctx, k_boundary_fun, dim3(grid_x_for(2 * maxn / std::max<long>(maxnni, 1) + maxnni, 256), n32),
                               dim3(256), 0, d_bnd);
            X2G_HIP(hipGetLastError());
        }
        if (d_tail) {
            X2G_LAUNCH(ctx, k_orb_tail, dim3(grid_x_for(maxnni, 128), (int)tails.size()), dim3(128), 0,
                               d_tail);
            X2G_HIP(hipGetLastError());
        }
    }
    return X2G_OK;
}

// ---------------------------------------------------------------------------
// Graph replay of x2g_relax_batch.  A call's launch sequence depends only on
// its signature: the items (arrays, parameters, scheme), their grids, the
// sweep counts, the sweep-plan settings and the workspace addresses.  The
// second call of a signature is captured into a HIP graph whose launch
// descriptors live in a device arena of the graph's own (the staging ring
// is reused by later calls, so a graph cannot point into it); later calls of
// the signature replay the graph.  The first call runs eagerly: it tunes the
// sweep plan, builds level schedules and sizes the workspace, so the capture
// makes no allocation and no synchronisation.  The captured sequence holds
// kernels only (no memset or copy nodes).
constexpr size_t kGraphArena = 1 << 20;   // first arena size; doubled on overflow
constexpr size_t kMaxGraphs = 64;         // cached graphs per context (LRU)

static std::string relax_signature(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_relax_item *items, int n,
                                   int maxsor1, int maxsor2)
{
    std::string k;
    auto put = [&](const auto &v) { k.append(reinterpret_cast<const char *>(&v), sizeof(v)); };
    put(n);
    put(maxsor1);
    put(maxsor2);
    put(g_plan_gen.load());
    put(ctx->ws[0].ptr);
    put(ctx->ws[3].ptr);
    for (int i = 0; i < n; ++i) {
        const x2g_relax_item &it = items[i];
        const x2g_grid *g = grids[it.grid_index];
        put(g);
        put(g->uid);
        put(it.fun_d);
        put(it.lhs_d);
        put(it.rhs_d);
        put(it.tail_fock1_d);
        put(it.rhs_f1_d);
        put(it.rhs_f2_d);
        put(it.rhs_f3_d);
        put(it.omega);
        put(it.isym);
        put(it.bkind);
        put(it.prec);
        put(it.method);
        put(it.ordering);
    }
    return k;
}

// Capture relax_batch_eager into a graph (the caller has checked the
// arguments).  On success the graph is cached under `key`; the captured work
// has not run yet.
static int relax_capture(x2g_ctx *ctx, const std::string &key, x2g_grid *const *grids,
                         const x2g_relax_item *items, int n, int maxsor1, int maxsor2,
                         x2g_ctx::GraphEntry **out)
{
    int rc = stage_flush(ctx);  // descriptors staged before the capture go the usual way
    if (rc) return rc;
    for (size_t cap = kGraphArena; cap <= 64 * kGraphArena; cap *= 2) {
        void *arena = nullptr;
        X2G_HIP(hipMalloc(&arena, cap));
        ctx->cap_arena = (char *)arena;
        ctx->cap_bytes = cap;
        ctx->cap_off = 0;
        ctx->cap_overflow = false;
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rc = relax_batch_eager(ctx, grids, items, n, maxsor1, maxsor2);
            e = hipStreamEndCapture(ctx->stream, &graph);
        }
        ctx->cap_arena = nullptr;
        const bool overflow = ctx->cap_overflow;
        if (e != hipSuccess || rc || !graph) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipFree(arena);
            (void)hipGetLastError();
            if (overflow) continue;
            set_error("x2g_relax_batch: graph capture failed (%s)", e != hipSuccess ? hipGetErrorString(e)
                                                                                      : "launch error");
            return rc ? rc : X2G_EHIP;
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e == hipSuccess)
            e = hipMemcpy(arena, ctx->cap_host.data(), ctx->cap_off, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (exec) (void)hipGraphExecDestroy(exec);
            (void)hipFree(arena);
            set_error("x2g_relax_batch: graph instantiation failed (%s)", hipGetErrorString(e));
            return X2G_EHIP;
        }
        if (ctx->graphs.size() >= kMaxGraphs) {
            // the least recently used graph may still run: wait for the
            // stream before releasing it and its arena
            e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) {
                (void)hipGraphExecDestroy(exec);
                (void)hipFree(arena);
                set_error("x2g_relax_batch: %s", hipGetErrorString(e));
                return X2G_EHIP;
            }
            auto victim = ctx->graphs.begin();
            for (auto it = ctx->graphs.begin(); it != ctx->graphs.end(); ++it)
                if (it->second.last_use < victim->second.last_use) victim = it;
            (void)hipGraphExecDestroy(victim->second.exec);
            (void)hipFree(victim->second.arena);
            ctx->graphs.erase(victim);
        }
        x2g_ctx::GraphEntry &g = ctx->graphs[key];
        g.exec = exec;
        g.arena = arena;
        g.ws0 = ctx->ws[0].ptr;
        g.ws3 = ctx->ws[3].ptr;
        ++ctx->graph_captures;
        *out = &g;
        return X2G_OK;
    }
    set_error("x2g_relax_batch: launch descriptors too large for a graph arena");
    return X2G_EINVAL;
}

extern "C" int x2g_relax_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_relax_item *items, int n,
                               int maxsor1, int maxsor2)
{
    X2G_CHECK(ctx && grids && (items || n == 0) && n >= 0, "x2g_relax_batch: null argument");
    // statistics time launches with events on the stream: eager only
    if (!ctx->graphs_on || ctx->stats_on || n == 0 || maxsor1 == 0)
        return relax_batch_eager(ctx, grids, items, n, maxsor1, maxsor2);
    for (int i = 0; i < n; ++i)
        X2G_CHECK(items[i].grid_index >= 0 && grids[items[i].grid_index], "x2g_relax_batch: bad grid index %d",
                  items[i].grid_index);
    const std::string key = relax_signature(ctx, grids, items, n, maxsor1, maxsor2);
    x2g_ctx::GraphEntry *ge = nullptr;
    auto it = ctx->graphs.find(key);
    if (it != ctx->graphs.end()) {
        ge = &it->second;
    } else {
        int &seen = ctx->graph_seen[key];
        if (seen < 0 || ++seen < 2) return relax_batch_eager(ctx, grids, items, n, maxsor1, maxsor2);
        if (ctx->graph_seen.size() > 4096) ctx->graph_seen.clear();
        int rc = relax_capture(ctx, key, grids, items, n, maxsor1, maxsor2, &ge);
        if (rc) {
            // the signature stays eager (the failed capture ran nothing)
            ctx->graph_seen[key] = -1;
            return relax_batch_eager(ctx, grids, items, n, maxsor1, maxsor2);
        }
    }
    ge->last_use = ++ctx->graph_clock;
    X2G_HIP(hipGraphLaunch(ge->exec, ctx->stream));
    ++ctx->graph_replays;
    return X2G_OK;
}

extern "C" int x2g_set_graphs(x2g_ctx *ctx, int on)
{
    X2G_CHECK(ctx, "x2g_set_graphs: null ctx");
    if (ctx->graphs_on && !on) {
        X2G_HIP(hipStreamSynchronize(ctx->stream));
        graphs_clear(ctx);
    }
    ctx->graphs_on = on != 0;
    return X2G_OK;
}

extern "C" int x2g_graph_counts(x2g_ctx *ctx, long long *captures, long long *replays)
{
    X2G_CHECK(ctx && captures && replays, "x2g_graph_counts: null argument");
    *captures = ctx->graph_captures;
    *replays = ctx->graph_replays;
    return X2G_OK;
}

extern "C" long long x2g_relax_workspace_bytes(int nni, int nmu, int prec)
{
    if (nni < 7 || nmu < 11 || (prec != 0 && prec != 1)) return -1;
    // the sizes x2g_relax_batch reserves per item, with the widest row
    // stride any sweep kind uses (nni+8 rounded up to 16 bytes, plus 16)
    const size_t elem = prec ? sizeof(float) : sizeof(double);
    const size_t stride = (size_t)(nni + 8 + 16 / elem - 1) / (16 / elem) * (16 / elem) + 16;
    const size_t pad = align256(stride * (nmu + 8) * elem);
    const size_t n = (size_t)nni * nmu, ns = (size_t)split_size(nni, nmu);
    if (prec == 0) return (long long)(2 * pad + align256(2 * ns * sizeof(double)));
    return (long long)(2 * pad + 2 * align256(std::max(n, ns) * sizeof(float)));
}

extern "C" int x2g_residual_norms(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_relax_item *items, int n,
                                  double *out_h)
{
    X2G_CHECK(ctx && grids && (items || n == 0) && out_h && n >= 0, "x2g_residual_norms: null argument");
    if (n == 0) return X2G_OK;
    long maxn = 0, maxnni = 0, maxnmu = 0;
    for (int i = 0; i < n; ++i) {
        X2G_CHECK(items[i].grid_index >= 0 && grids[items[i].grid_index], "x2g_residual_norms: bad grid index");
        const x2g_grid *g = grids[items[i].grid_index];
        maxn = std::max<long>(maxn, (long)g->n());
        maxnni = std::max<long>(maxnni, g->nni);
        maxnmu = std::max<long>(maxnmu, g->nmu);
    }
    // partial-sum slots per problem: one per wave of the residual launch
    const long nslot = maxnni <= kResThreads ? (maxnmu - 5 + kResRows - 1) / kResRows * ((maxnni + 63) / 64)
                                             : (maxn + 255) / 256 * 4;
    int rc = ws_reserve(ctx, 1, (size_t)(2 * nslot * n + 2 * n) * sizeof(double));
    if (rc) return rc;
    double *part = (double *)ctx->ws[1].ptr, *acc = part + 2 * nslot * n;
    std::vector<ResDesc> res;
    for (int i = 0; i < n; ++i) {
        const x2g_relax_item &it = items[i];
        X2G_CHECK(it.grid_index >= 0 && grids[it.grid_index], "x2g_residual_norms: bad grid index");
        X2G_CHECK(it.fun_d && it.lhs_d && it.rhs_d, "x2g_residual_norms: item %d has a null array", i);
        x2g_grid *g = grids[it.grid_index];
        ResDesc r{};
        r.fun = it.fun_d;
        r.isym = it.isym;
        r.lhs = it.lhs_d;
        r.rhs = it.rhs_d;
        r.rf[0] = it.rhs_f1_d;
        r.rf[1] = it.rhs_f1_d ? it.rhs_f2_d : nullptr;
        r.rf[2] = it.rhs_f1_d ? it.rhs_f3_d : nullptr;
        r.bmu = it.bkind == 0 ? g->bmu_orb : g->bmu_pot;
        r.dnu = g->dnu;
        r.res_f = nullptr;
        r.part = part + 2 * nslot * i;
        for (int q = 0; q < 16; ++q) r.c[q] = g->coef[q];
        r.nni = g->nni;
        r.nmu = g->nmu;
        res.push_back(r);
    }
    size_t doff = 0;
    const ResDesc *d_res;
    if ((rc = stage_descs(ctx, res, doff, &d_res))) return rc;
    // slots of strips beyond a smaller problem's rows are not written
    if ((rc = x2g_fill(ctx, part, 2 * (size_t)nslot * n, 0.0))) return rc;
    bool prod = false;
    for (auto &r : res) prod = prod || r.rf[0];
    if ((rc = launch_residual(ctx, d_res, n, maxnni, maxnmu, maxn, prod))) return rc;
    X2G_LAUNCH(ctx, k_sum_parts, dim3(n), dim3(64), 0, part, (int)nslot, acc);
    X2G_HIP(hipGetLastError());
    X2G_HIP(hipMemcpyAsync(out_h, acc, 2 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    return X2G_OK;
}
