/*
 * x2gpu -- MI355X (gfx950) relaxation engine for finite-difference diatomic
 * Hartree-Fock (the x2dhf method).  Plain C ABI: every entry point takes plain
 * pointers and sizes, returns an int status (X2G_OK = 0) and never throws.
 * All array pointers named *_d are device pointers in HBM.
 *
 * Storage conventions (identical to the reference):
 *   - a function f(nu,mu) on an nni x nmu grid is column-major, nu fastest:
 *     f[(mu-1)*nni + nu-1]  (unpadded, "nni*nmu" arrays);
 *   - a relaxation work array is padded to (nni+8) x (nmu+8) with
 *     W(nu+4,mu+4) = f(nu,mu) (inout.F90:30-36).
 *
 * Reference interfaces replaced (paths relative to the x2dhf source tree):
 *   x2g_mcsor          <- mcsor(isym,fun,lhs,rhs,b,d,indx1,..)   src/sormcsor.F90:143
 *                         and mcsor_pthread_(isym,nthreads,wk2,lhs,rhs)
 *                                                               src/c/mcsor_pthread.c:32
 *   x2g_mcsor_host     <- the same, with host arrays (copies in and out)
 *   x2g_relax_ordered  <- sor / mcsor in any mesh ordering    src/sormcsor.F90:20,143
 *                         (mesh, src/initCBAllocArrays.F90:2408-2540)
 *   x2g_putin/putout   <- putin/putout, putinc_/putoutc_         src/inout.F90:23,165,195,266
 *   x2g_relax_batch    <- coulExch_pthread_() + relax_single_pot_pthread_
 *                                                               src/c/coulExch_pthread.c:19
 *                         and the macro-iteration loops of orbSOR/orbMCSOR,
 *                         coulExchSOR/coulExchMCSOR            src/relaxOrbs.F90:92-139,
 *                                                               src/relaxPots.F90:97-104
 *   x2g_lincomb        <- the daxpy/prod chains of fockHF       src/fock.F90:88-184
 *   x2g_wdot           <- ddot(wgt,.) quadratures of norm/ortho/EaHF/coulij/exchij
 *                         src/normOrtho.F90:41-50, src/integrals.F90:11-30,372-399
 *   x2g_oneel          <- oneelij / oneelii                    src/integrals.F90:406,37
 *   x2g_moments        <- coulMom / exchMom                    src/scfUtils.F90:490,615
 *   x2g_pot_tail       <- coulAsympt / exchAsympt (vcoul/vexch) src/scfUtils.F90:275,310
 *   x2g_setci          <- setCiOrb                             src/scfUtils.F90:380
 *   x2g_guess_orbital  <- initHyd (hydrogenic LCAO start)      src/initOrbitalsPotentials.F90:714-880
 *   x2g_guess_coulomb  <- initCoulomb (Thomas-Fermi start)     src/coulombExchangePotentials.f90:14-71
 *   x2g_guess_exchange <- initExchange (1/r start)             src/coulombExchangePotentials.f90:74-147
 *
 * Batched forms (many problems, possibly on different grids, per launch; the
 * scalars stay on the device until x2g_gather):
 *   x2g_lincomb_batch  <- fockHF and the relaxPots right-hand sides for all
 *                         orbitals of several jobs  src/fock.F90:88-184, src/relaxPots.F90:62
 *   x2g_normortho_batch<- norm + setCiOrb + ortho of one orbital of each job
 *                                                  src/normOrtho.F90:14-59,103-190
 *   x2g_eahf_batch     <- EaHF (diagonal Lagrange multipliers)
 *                                                  src/lagrangeMultipliers.F90:21-107
 *   x2g_moments_batch  <- mpoleMoments (coulMom/exchMom of every orbital/pair)
 *                                                  src/scfUtils.F90:446-850
 *   x2g_pot_tail_batch <- potAsympt (coulAsympt/exchAsympt of every potential)
 *                                                  src/scfUtils.F90:209-329
 */
#ifndef X2GPU_H
#define X2GPU_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define X2G_OK 0
#define X2G_EINVAL (-1)  /* bad argument (size, pointer, unsupported grid) */
#define X2G_EHIP (-2)    /* HIP runtime error */
#define X2G_ENOMEM (-3)  /* device allocation failed */

typedef struct x2g_ctx x2g_ctx;   /* device + stream + workspace */
typedef struct x2g_grid x2g_grid; /* one (nni,nmu) grid and its coefficient arrays */

/* ---- library / context ------------------------------------------------- */
int x2g_version(void);
const char *x2g_last_error(void);
/* Create a context on HIP device `device` with its own non-blocking stream. */
int x2g_ctx_create(int device, x2g_ctx **out);
int x2g_ctx_destroy(x2g_ctx *ctx);
int x2g_ctx_sync(x2g_ctx *ctx);
/* Return the context's hipStream_t (as void*). */
void *x2g_ctx_stream(x2g_ctx *ctx);
/* Free and total device memory of the context's GPU (bytes). */
int x2g_mem_info(x2g_ctx *ctx, size_t *free_bytes, size_t *total_bytes);
/* Time the sweep launches of this context with HIP events on its stream: one
 * event pair around each run of back-to-back sweeps (a relaxation's maxsor2
 * sweeps), counted as that many launches, so the average launch time includes
 * the launch-to-launch gaps (on != 0 starts and clears the counters, 0 stops). */
int x2g_stats_enable(x2g_ctx *ctx, int on);
/* out[6] = {launches, milliseconds, algorithmic bytes} of the FP64 sweeps,
 * then of the FP32 sweeps.  Algorithmic bytes per relaxed point and sweep:
 * 4 words (read u, write u, read lhs, read rhs). */
int x2g_stats_get(x2g_ctx *ctx, double *out);

/* ---- device memory ------------------------------------------------------ */
/* Blocks freed with x2g_free stay with the context and are handed out again
 * by x2g_malloc for the same size (no synchronisation: all work of a
 * context is ordered by its stream); x2g_ctx_destroy releases them, and
 * x2g_mem_info counts them as free. */
int x2g_malloc(x2g_ctx *ctx, size_t bytes, void **dptr);
int x2g_free(x2g_ctx *ctx, void *dptr);
int x2g_h2d(x2g_ctx *ctx, void *dst_d, const void *src_h, size_t bytes);  /* synchronous */
int x2g_d2h(x2g_ctx *ctx, void *dst_h, const void *src_d, size_t bytes);  /* synchronous */
int x2g_d2d(x2g_ctx *ctx, void *dst_d, const void *src_d, size_t bytes);  /* stream-ordered */
/* dst[i] = value for i < n (fill kernel, never a memset node). */
int x2g_fill(x2g_ctx *ctx, double *dst_d, size_t n, double value);

/* ---- grid --------------------------------------------------------------- */
/* Upload one grid.  coef = {dmu1[4], dmu2[4], dni1[4], dni2[4], exeven[5],
 * diag, hni, hmu, r} (25 doubles, initSuppl initCBAllocArrays.F90:1979-2041);
 * suppl = the 14 nni*nmu coefficient arrays in the reference's i4b order
 * (borb, bpot, d, e, f0, f1, f2, f3, f4, g, wjac1, wjac2, wgt1, wgt2);
 * vec = {vxi[nmu], vxisq[nmu], veta[nni], vetasq[nni]}. */
int x2g_grid_create(x2g_ctx *ctx, int nni, int nmu, const double *coef, const double *suppl,
                    const double *vec, x2g_grid **out);
int x2g_grid_destroy(x2g_grid *g);

/* ---- starting functions of a job ---------------------------------------- */
/* LCAO start of one orbital, c1 phi_A(n1,l1,m; za) + c2 phi_B(n2,l2,m; zb),
 * with phi the hydrogenic function in the reference's normalisation and sign
 * (unnormalised sum, as initHyd leaves it); c1 or c2 = 0 omits that centre.
 * Replaces initHyd / hydrogenOrbA / hydrogenOrbB
 * src/initOrbitalsPotentials.F90:714-880. */
typedef struct {
    int m;               /* magnetic quantum number of the orbital */
    int n1, l1, n2, l2;  /* hydrogenic quantum numbers on centres A and B */
    double c1, za;       /* mixing coefficient and exponent on centre A */
    double c2, zb;       /* the same on centre B */
} x2g_lcao;
int x2g_guess_orbital(x2g_grid *g, double *orb_d, const x2g_lcao *lc);
/* Thomas-Fermi start of a Coulomb potential (times f4): -|c1| V_TF(r_A)
 * [- |c2| V_TF(r_B) when two_centres != 0].  Replaces initCoulomb + pottf
 * src/coulombExchangePotentials.f90:14-71,222-239. */
int x2g_guess_coulomb(x2g_grid *g, double *pot_d, double c1, double c2, int two_centres);
/* 1/r start of an exchange potential (times f4): c1/r_A + c2/r_B on the
 * interior nu = 2..nni-1, mu >= 2.  Replaces initExchange
 * src/coulombExchangePotentials.f90:74-147. */
int x2g_guess_exchange(x2g_grid *g, double *pot_d, double c1, double c2);
/* Device pointer of supplementary array k (0..13, i4b order) of grid g. */
const double *x2g_grid_suppl(x2g_grid *g, int k);

/* ---- relaxation ---------------------------------------------------------- */
/* Sweep engine for all later calls of this process: 0 = default: the
 * pipelined strip sweep with a fixed lane map and residue-split lhs/rhs where
 * it fits, else 2, else 1; 1 = one launch per colour pass; 2 = the first
 * pipelined strip kernel (natural lhs/rhs; falls back to 1 for grids too wide
 * for its LDS ring).  strip_rows > 0 fixes the strip height (rounded up to a
 * multiple of 4), <= 0 lets the library choose. */
int x2g_set_sweep_engine(int engine, int strip_rows);

/* Schedule of ordered sweeps (x2g_relax_ordered, relax items with a method
 * other than the col-wise MCSOR) for all later calls of this process:
 * 1 = default: the sweeps pipelined, each a fixed number of dependency
 * levels behind the one before, one barrier per step of all of them;
 * 0 = one dependency level of one sweep per barrier (X2G_LEVELS=phase).
 * Both give the sequential routine's result bit for bit (sormcsor.F90:20-132,
 * 143-386 with one thread).  No counterpart in the reference: a choice of
 * this library's schedule. */
int x2g_set_levels_pipeline(int on);

/* Sweep autotuner (on by default; X2G_AUTOTUNE=0 turns it off): on the first
 * relaxation of a batch shape (precision, ring class of the widest grid,
 * number of ring classes, total rows rounded to a power of sqrt 2) the
 * candidate k_sweep plans -- blocks of 4 or 2 rows, workgroups per CU 1, 2, 3
 * or as many as the LDS ring allows, one launch or one per ring class -- are
 * timed on that batch and the fastest is kept for the shape.  on = 0 / 1
 * turns it off / on, -1 forgets the tuned shapes.  No reference counterpart. */
int x2g_set_autotune(int on);
/* Fix the k_sweep plan for later calls of this process: block_rows 4 or 2
 * (0: 4), wg_per_cu workgroups per CU the segment planner fills (0: as many
 * as the LDS ring allows), split 1 = one launch per ring class.  Any nonzero
 * value keeps the autotuner out; (0, 0, 0) hands the choice back to it.  The
 * results do not depend on the plan.  No reference counterpart. */
int x2g_set_sweep_plan(int block_rows, int wg_per_cu, int split);
/* Replay x2g_relax_batch from captured HIP graphs on this context (default
 * off; X2G_GRAPHS=1 at context creation turns it on).  The second call with
 * the same signature -- items, grids, sweep counts, sweep-plan settings and
 * workspace -- is captured into a graph of its kernel launches, with the
 * launch descriptors in a device arena of the graph's own; later calls with
 * that signature launch the graph.  Results are those of the eager calls.
 * Not while sweep statistics are on.  on = 0 releases the graphs.  No
 * reference counterpart (the reference keeps its worker threads across
 * calls, c/mcsor_tpool.c:41-199; this keeps the launch sequence). */
int x2g_set_graphs(x2g_ctx *ctx, int on);
/* Graphs captured and replays launched on this context so far. */
int x2g_graph_counts(x2g_ctx *ctx, long long *captures, long long *replays);
/* The tuned shapes as text, one line per shape: "key: br B wpc W split S
 * (batches N); br/wpc/split time ..." (times per sweep of the candidates).
 * Writes at most len bytes including the terminating 0; returns the full
 * length. */
int x2g_tune_report(char *buf, int len);

/* putin (inout.F90:23): embed fun_d (nni*nmu) into the padded work_d and fill
 * the mirror halos by parity isym (+1 even, -1 odd). */
int x2g_putin(x2g_grid *g, int isym, const double *fun_d, double *work_d);
/* putout (inout.F90:165): copy the interior of work_d back into fun_d. */
int x2g_putout(x2g_grid *g, double *fun_d, const double *work_d);

/* MCSOR drop-in (sormcsor.F90:143): nsweeps five-colour sweeps of the
 * 17-point cross stencil on the padded work_d, each followed by the boundary
 * extrapolation of the mu=1, nu=1 and nu=nni lines.  bkind selects b(mu):
 * 0 = borb (orbitals), 1 = bpot (potentials).  work_d is updated in place. */
int x2g_mcsor(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d,
              const double *rhs_d, double omega, int nsweeps);
/* x2g_mcsor without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and result as x2g_mcsor, with the call's workspace in a static
 * device array of 24 MiB per GPU (held until the process ends; enough for a
 * grid of about 560 x 1000 points), instead of the context's growable
 * workspace.  Larger grids are rejected with X2G_EINVAL.  No hipMalloc /
 * hipFree, and sweep timing (x2g_stats_enable) is not recorded for it.  Every
 * call on a GPU shares that array: calls that overlap on one GPU (two
 * contexts or two host threads) race; calls on one context are ordered by
 * its stream. */
int x2g_mcsor_noalloc(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d,
                      const double *rhs_d, double omega, int nsweeps);
/* sor (sormcsor.F90:20-132, method X2G_SOR) or mcsor (sormcsor.F90:143-386,
 * method X2G_MCSOR, one thread) in mesh ordering `ordering`
 * (initCBAllocArrays.F90:2408-2810; the reference's default is middle) on the
 * padded work_d, in place: nsweeps sweeps, each followed by the boundary
 * lines.  The points are updated in exactly the reference's sequential order
 * of dependencies: each sweep runs as a list of phases of mutually
 * independent points (for sor: the anti-diagonals of the ordering's
 * (position in row, row) plane), all in one workgroup.  The orderings are
 * the reference's `order` card values (read_order, inputInterface.F90:2397):
 * col-wise, middle, middlemt, row-wise, rrow-wis, test1, test2, test3,
 * test4, middle2.  Method X2G_SOR_BACKWARD is sor walking the ordering from
 * its last point to its first: the backward sweeps of the reference's
 * altsweeps option (isstart = ngrid1, isstop = 1, isstep = -1 for odd
 * orbitals, doSCF.F90:403-411, read by sor at sormcsor.F90:42). */
#define X2G_MCSOR 0
#define X2G_SOR 1
#define X2G_SOR_BACKWARD 2
#define X2G_ORDER_COLWISE 0
#define X2G_ORDER_MIDDLE 1
#define X2G_ORDER_MIDDLEMT 2
#define X2G_ORDER_ROWWISE 3
#define X2G_ORDER_RROWWISE 4
#define X2G_ORDER_TEST1 5
#define X2G_ORDER_TEST2 6
#define X2G_ORDER_TEST3 7
#define X2G_ORDER_TEST4 8
#define X2G_ORDER_MIDDLE2 9
int x2g_relax_ordered(x2g_grid *g, int isym, int bkind, double *work_d, const double *lhs_d,
                      const double *rhs_d, double omega, int nsweeps, int method, int ordering);
/* Host-pointer drop-in for mcsor/sor: work (padded), lhs, rhs on the host;
 * copies in, relaxes on the device, copies work back.  PCIe-inclusive. */
int x2g_mcsor_host(x2g_grid *g, int isym, int bkind, double *work, const double *lhs,
                   const double *rhs, double omega, int nsweeps);

/* One relaxation problem: maxsor1 macro-iterations of
 *   [orbital tail factors] putin -> maxsor2 sweeps -> putout [orbital tail]
 * on fun_d (relaxOrbs.F90:92-139 / relaxPots.F90:97-104).
 * tail_fock1_d != NULL marks an orbital: its tail rows mu=nmu-3..nmu are
 * recomputed after each macro-iteration from the decay factors of
 * fock1/f1 (orbAsymptGet/Set, relaxOrbs.F90:104-124).
 * prec: 0 = FP64 sweeps; 1 = FP32 correction sweeps on A e = rhs - A u with
 * the residual and the update u += e formed in FP64 (mixed precision). */
typedef struct {
    double *fun_d;
    const double *lhs_d;
    const double *rhs_d;
    const double *tail_fock1_d;
    double omega;
    int isym;
    int bkind;
    int prec;
    int grid_index; /* index into the grids[] array of x2g_relax_batch */
    /* Optional right-hand side as a product of up to three arrays (the
     * potentials' psi_i psi_j g of relaxPots.F90:97-104): when rhs_f1_d is
     * not NULL, rhs = rhs_f1 * rhs_f2 * rhs_f3 (a NULL rhs_f2 / rhs_f3 counts
     * as 1), formed where the relaxation reads it (the FP64 residual of a
     * prec = 1 item), and rhs_d is scratch that the call may overwrite with
     * the product.  Zero-initialise all three to pass rhs_d alone. */
    const double *rhs_f1_d;
    const double *rhs_f2_d;
    const double *rhs_f3_d;
    /* Relaxation scheme (prec = 0 items): method X2G_MCSOR (five passes),
     * X2G_SOR or X2G_SOR_BACKWARD, over the points in mesh ordering X2G_ORDER_* (one of the
     * reference's ten, initCBAllocArrays.F90:2408-2810).  Zero (MCSOR,
     * col-wise) is the pipelined colour sweep; any other pair runs the
     * reference's sequential point order exactly (x2g_relax_ordered). */
    int method;
    int ordering;
} x2g_relax_item;

/* Relax n problems, possibly on different grids, together: every sweep of
 * every problem is issued in one fused launch (items of another method or
 * ordering: all their sweeps in one launch, a workgroup per item). */
int x2g_relax_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_relax_item *items, int n,
                    int maxsor1, int maxsor2);

/* Device workspace bytes (an upper bound) that one relaxation problem on an
 * nni x nmu grid of precision prec adds to the context's workspace while
 * x2g_relax_batch runs (padded sweep buffers and the colour-ordered
 * lhs/rhs pairs); the workspace is shared by the items of a call and kept
 * between calls.  No device call; -1 for bad sizes.  No reference
 * counterpart (the reference's per-thread temporaries,
 * c/relax_single_pot_pthread.c:65-67, are host mallocs). */
long long x2g_relax_workspace_bytes(int nni, int nmu, int prec);

/* FP64 residual norm^2 sum (rhs - A u)^2 over the relaxed interior of each
 * item, and sum rhs^2 (for relative residuals).  out[2*i], out[2*i+1]. */
int x2g_residual_norms(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_relax_item *items, int n,
                       double *out_h);

/* ---- elementwise / quadrature ------------------------------------------- */
/* out = base_scale*base + sum_k c_k * a_k * b_k * cc_k (b_k, cc_k may be NULL;
 * base may be NULL).  out may alias any input (pure elementwise). */
typedef struct {
    double c;
    const double *a_d;
    const double *b_d;
    const double *cc_d;
} x2g_term;
int x2g_lincomb(x2g_grid *g, double *out_d, const double *base_d, double base_scale,
                const x2g_term *terms, int nterms);

/* out_h[k] = sum_i w_k[i] * a_k[i] * b_k[i] * cc_k[i] over nni*nmu points for
 * k < nterms (b_k, cc_k may be NULL); deterministic two-pass FP64 reduction. */
typedef struct {
    const double *w_d;
    const double *a_d;
    const double *b_d;
    const double *cc_d;
} x2g_dot;
int x2g_wdot(x2g_grid *g, const x2g_dot *terms, int nterms, double *out_h);

/* oneelij (integrals.F90:406): <psi2 | T + m^2 e + f0 | psi1> with wgt1, where T
 * is the 9-point (nu,mu) Laplacian of psi1 (diffnu/diffmu, utils.f90:599-695)
 * on the mirror-padded psi1 (parity isym). out_h[0] = total, out_h[1] = the
 * kinetic part <psi2|T + m^2 e|psi1>, out_h[2] = <psi2|f0|psi1>. */
int x2g_oneel(x2g_grid *g, const double *psi1_d, const double *psi2_d, int isym, int m,
              double *out_h);

/* Multipole moments (coulMom/exchMom, scfUtils.F90:490-850):
 * out_h[k-1] = sum wgt2*f4*psi1*psi2 * N_kq P_k^q(cos th) r^k, k = 1..mpole,
 * q = dm (0 for Coulomb moments). */
int x2g_moments(x2g_grid *g, const double *psi1_d, const double *psi2_d, int dm, int mpole,
                double *out_h);

/* Dirichlet tail mu=nmu-3..nmu of a potential from its multipole moments
 * (coulAsympt/exchAsympt, scfUtils.F90:275-329).  kind 0: Coulomb
 * (mom = cmulti of the orbital, 1/r monopole included); kind 1: exchange with
 * |Delta m| = dm (mom = exc(di,qu,oc,he,5..8)). */
int x2g_pot_tail(x2g_grid *g, double *pot_d, int kind, int dm, int mpole, const double *mom);

/* setCiOrb (scfUtils.F90:380-406): impose g (ihsym=+1) or u (-1) symmetry. */
int x2g_setci(x2g_grid *g, double *psi_d, int ihsym);

/* ---- batched operations -------------------------------------------------- */
/* out = base_scale*base + sum of terms[first_term .. first_term+nterms-1], on
 * grids[grid_index]; all ops of a call run in one launch, so no op may read
 * another op's output. */
typedef struct {
    double *out_d;
    const double *base_d; /* may be NULL */
    double base_scale;
    int first_term, nterms;
    int grid_index;
} x2g_lincomb_op;
int x2g_lincomb_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_lincomb_op *ops, int nops,
                      const x2g_term *terms, int nterms_total);

/* norm + setCiOrb + ortho of one orbital per item (normOrtho.F90:14-59,
 * 103-190), with the quadrature weights wgt2*f4 of the item's grid:
 *   s = <psi|psi> (written to norm_out_d when not NULL);
 *   psi <- setci(psi) (ihomo = +1/-1; 0 = no symmetry);
 *   ovla_k = <psi|other_k>/sqrt(s);  a = sum ovla_k^2;
 *   psi <- sqrt(1+a)/(1-a^2) * (psi/sqrt(s) - sum_k ovla_k other_k)
 * (the reference's sequence folded into one update; nothers = 0 leaves the
 * plain normalisation).  normalise = 0: only the symmetry is imposed. */
#define X2G_MAX_ORTHO 16
typedef struct {
    double *psi_d;
    const double *others_d[X2G_MAX_ORTHO];
    int nothers;
    int ihomo;
    int normalise;
    int grid_index;
    double *norm_out_d;
} x2g_normortho_item;
int x2g_normortho_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_normortho_item *items, int n);

/* ee = <psi|h|psi> + <psi|cpot psi> - <psi|xpot> (EaHF, lagrangeMultipliers.F90:
 * 21-107; h as x2g_oneel) written to ee_out_d, one item per orbital. */
typedef struct {
    const double *psi_d, *cpot_d, *xpot_d;
    int isym, m, grid_index;
    double *ee_out_d;
} x2g_eahf_item;
int x2g_eahf_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_eahf_item *items, int n);

/* Multipole moments (as x2g_moments) of n pairs; out_h[8*k + q] (synchronises). */
typedef struct {
    const double *psi1_d, *psi2_d;
    int dm, grid_index;
} x2g_moment_item;
int x2g_moments_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_moment_item *items, int n, int mpole,
                      double *out_h);

/* Potential tails (as x2g_pot_tail) of n potentials in one launch. */
typedef struct {
    double *pot_d;
    int kind, dm, grid_index;
    double mom[8];
} x2g_tail_item;
int x2g_pot_tail_batch(x2g_ctx *ctx, x2g_grid *const *grids, const x2g_tail_item *items, int n, int mpole);

/* Copy counts[k] doubles from each src_d[k] into out_h, packed in order, with
 * one device-to-host transfer (synchronises). */
int x2g_gather(x2g_ctx *ctx, const double *const *src_d, const int *counts, int n, double *out_h);

/* ---- SCF control on the device ------------------------------------------ */
/* The end-of-iteration bookkeeping of the SCF loop (doSCF.F90:546-744: the
 * orbital-energy changes, the energy / norm threshold tests with their
 * nscfextra confirmations and resets, the saturation test, the iteration
 * limit, the release of the frozen starting orbitals, the multipole-refresh
 * decision) plus the failure guards (a NaN norm or energy, detect.F90:9-58;
 * an energy change above orbEnergyIncThld, doSCF.F90:612-616) and the FP64
 * finish of a mixed-precision job, for one job per record.  The record
 * lives in device memory and is the job's control state between iterations;
 * the caller writes it once (settings and the state after the set-up pass)
 * and reads it back to learn the decisions.  Orbital i is index i-1. */
#define X2G_CTRL_MAXO 16
typedef struct {
    /* settings */
    double ethld;                  /* 10^-ienterm: orbital-energy threshold */
    double nthld;                  /* 10^-inoterm: orbital-norm threshold */
    double thren4init;             /* 10^-ienterm4init: frees the frozen starting orbitals */
    double recalc_mm;              /* multipole refresh when |eeconv / max d eps| exceeds it */
    double inc_thld;               /* orbEnergyIncThld: a larger energy change is divergence */
    /* state */
    double ee[X2G_CTRL_MAXO];      /* orbital energies (diagonal Lagrange multipliers) */
    double eeold[X2G_CTRL_MAXO];
    double deltaee[X2G_CTRL_MAXO];
    double orbnorm[X2G_CTRL_MAXO]; /* <psi|psi> before normalisation */
    double eeconv, eeconvprev, dnmaxprev, deltaeemax, deltanormmax;
    int norb, maxscf, nscf2skip, nenlast, nnolast, nscfextra, maxresets, polish;
    int ifixorb[X2G_CTRL_MAXO];    /* 1: frozen (a hydrogenic start until the potentials settle) */
    int itouch[X2G_CTRL_MAXO];
    int mm_touch[X2G_CTRL_MAXO];   /* orbitals whose moments the refresh recomputes */
    int iscf, nresets, nresets4no, norbenergythld, norbnormthld, noenergydec, nonormdec;
    int lenergythld, lnormthld, lfixorb, lfixorb4init;
    int mm_due;                    /* out: refresh the multipole moments now */
    int prec;                      /* 1: FP32 correction sweeps; set to 0 by the FP64 finish */
    int polish_end;                /* no threshold stop before this iteration (FP64 finish) */
    int stopcode;                  /* 1 energy, 2 norm, 3 limit, 4 saturation, 5 resets, 6 diverged, 7 NaN */
    int stop;                      /* out: the job stops after this iteration */
} x2g_scf_ctrl;

/* One SCF iteration's control of n jobs in one launch: for job k, with
 * rec_d[k][0..norb-1] the norms and rec_d[k][16..16+norb-1] the orbital
 * energies this iteration produced (x2g_normortho_batch norm_out_d,
 * x2g_eahf_batch out_d), advance ctrl_d[k] by one iteration.  The records
 * stay on the device; read them with x2g_gather (sizeof(x2g_scf_ctrl) / 8
 * doubles each).  No reference counterpart as an interface: the reference
 * runs these rules on the host inside scf (doSCF.F90:546-744). */
int x2g_scf_control(x2g_ctx *ctx, x2g_scf_ctrl *const *ctrl_d, const double *const *rec_d, int n);

/* sizeof(x2g_scf_ctrl), for callers that mirror the record in another
 * language (the Fortran driver binds it with ISO_C_BINDING). */
int x2g_scf_ctrl_bytes(void);

#ifdef __cplusplus
}
#endif
#endif
