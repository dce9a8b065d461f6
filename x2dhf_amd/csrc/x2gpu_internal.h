// Internal declarations of the x2gpu library (MI355X / gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>
#include "../../include/x2gpu.h"

namespace x2g {

void set_error(const char *fmt, ...);

#define X2G_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::x2g::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                 \
                             hipGetErrorString(e_));                                    \
            return X2G_EHIP;                                                            \
        }                                                                               \
    } while (0)

#define X2G_CHECK(cond, ...)                                                            \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::x2g::set_error(__VA_ARGS__);                                              \
            return X2G_EINVAL;                                                          \
        }                                                                               \
    } while (0)

// Growable device workspace owned by a context (never freed while a launch
// that uses it may still run: growth synchronises the stream first).
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace x2g

struct x2g_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    size_t max_lds = 0;
    x2g::Workspace ws[6];          // 0: padded work buffers, 1: reduction partials,
                                   // 2: launch descriptors, 3: FP32 staging,
                                   // 4: padded buffers of batched quadratures,
                                   // 5: batched reduction partials
    void *pinned = nullptr;        // pinned host staging for descriptors/results
    size_t pinned_bytes = 0;
    void *pinned_out = nullptr;    // pinned host buffer for gathered results
    size_t pinned_out_bytes = 0;
    size_t stage_off = 0;          // ring offset into ws[2] / pinned
    size_t stage_lo = 0, stage_hi = 0;  // staged bytes not yet copied to ws[2]: [lo, hi)
    // sweep-kernel timing (x2g_stats_enable): an event pair around each run of
    // back-to-back sweep launches, resolved at x2g_stats_get or when the
    // pending list is long
    bool stats_on = false;
    std::vector<hipEvent_t> ev_pool;
    struct Pending { hipEvent_t a, b; double bytes; int prec, launches; };
    std::vector<Pending> pending;
    double st_launch[2] = {0, 0}, st_ms[2] = {0, 0}, st_bytes[2] = {0, 0};
    // device blocks of x2g_malloc: sizes of the live ones, and freed ones
    // kept for reuse by size (all work on them is ordered by `stream`)
    std::unordered_map<void *, size_t> live;
    std::multimap<size_t, void *> cache;
    size_t cached_bytes = 0;
    // device address of the static scratch array of x2g_mcsor_noalloc on
    // this context's device (shared by every context on that device)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // x2g_relax_batch replayed from captured HIP graphs (x2g_set_graphs):
    // the instantiated graph of each call signature with the device arena
    // that holds its launch descriptors, and the signatures seen once (a
    // signature is captured on its second call, after the first has tuned
    // the sweep plan and sized the workspace)
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        void *arena = nullptr;
        void *ws0 = nullptr, *ws3 = nullptr;  // workspace the graph addresses
        long last_use = 0;
    };
    bool graphs_on = false;
    std::unordered_map<std::string, GraphEntry> graphs;
    std::unordered_map<std::string, int> graph_seen;
    long graph_clock = 0, graph_captures = 0, graph_replays = 0;
    // during a capture: descriptors are staged into cap_host and addressed in
    // cap_arena, where they are copied once the capture has ended
    char *cap_arena = nullptr;
    size_t cap_bytes = 0, cap_off = 0;
    bool cap_overflow = false;     // a capture's descriptors did not fit its arena
    std::vector<char> cap_host;
};

// Level schedule of an ordered (MC)SOR sweep on one grid (device arrays):
// phases off[k] .. off[k+1]-1 of pts (unpadded point indices) are mutually
// independent and run in order.  The pipelined form: one sweep's point
// updates and boundary extrapolations on one level scale (levels poff[l] ..
// poff[l+1]-1 of pops), consecutive sweeps run lag levels apart.
struct x2g_levels {
    int *off = nullptr;
    int *pts = nullptr;
    int nphase = 0;
    int *poff = nullptr;
    int *pops = nullptr;
    int nlev = 0, lag = 0;
};

struct x2g_grid {
    x2g_ctx *ctx = nullptr;
    unsigned long long uid = 0;    // unique per grid created in this process
    std::map<int, x2g_levels> levels;  // by (method, ordering), built on first use
    int nni = 0, nmu = 0;
    double coef[25];               // dmu1,dmu2,dni1,dni2 (16), exeven (5), diag, hni, hmu, r
    double *suppl = nullptr;       // 14 * nni*nmu
    double *bmu_orb = nullptr;     // b(mu) for orbitals  (borb(1,mu))
    double *bmu_pot = nullptr;     // b(mu) for potentials (bpot(1,mu))
    double *dnu = nullptr;         // d(nu)  (d(nu,1))
    float *bmu_orb_f = nullptr, *bmu_pot_f = nullptr, *dnu_f = nullptr;
    double *vxi = nullptr, *vxisq = nullptr, *veta = nullptr, *vetasq = nullptr;
    size_t n() const { return (size_t)nni * nmu; }
    size_t npad() const { return (size_t)(nni + 8) * (nmu + 8); }
};

namespace x2g {
// Reserve ws[k] of at least `bytes` (synchronises the stream on growth).
int ws_reserve(x2g_ctx *ctx, int k, size_t bytes);
int pinned_reserve(x2g_ctx *ctx, size_t bytes);
// Copy `bytes` of launch descriptors to device memory through the context's
// staging ring (pinned host buffer -> stream-ordered copy, issued by the next
// stage_flush).  A region is reused only after the stream has been
// synchronised past its last use.
int stage_bytes(x2g_ctx *ctx, const void *src, size_t bytes, const void **dev_out);
// Issue one stream-ordered copy of everything staged since the last flush.
int stage_flush(x2g_ctx *ctx);
// Release every captured graph of the context.
void graphs_clear(x2g_ctx *ctx);
// Launch a kernel on the context's stream once the staged descriptors are on
// their way (every launch goes through here, so a call's descriptors cost one
// copy however many arrays it stages).
#define X2G_LAUNCH(ctx, kernel, grid, block, lds, ...)                                  \
    do {                                                                                \
        const int rc_ = ::x2g::stage_flush(ctx);                                        \
        if (rc_) return rc_;                                                            \
        hipLaunchKernelGGL(kernel, grid, block, lds, (ctx)->stream, __VA_ARGS__);       \
    } while (0)
template <typename D>
inline int stage_n(x2g_ctx *ctx, const D *v, size_t n, const D **out)
{
    const void *p = nullptr;
    int rc = stage_bytes(ctx, v, n * sizeof(D), &p);
    *out = reinterpret_cast<const D *>(p);
    return rc;
}

// Global-memory view of a device pointer: loads through it are global (not
// flat) loads, which only the vector-memory counter tracks, so waiting for a
// scalar load (a term's pointers) or an LDS result does not also wait for them.
typedef const double __attribute__((address_space(1))) *gdptr;
typedef double __attribute__((address_space(1))) *gdptr_w;
typedef const float __attribute__((address_space(1))) *gfptr;
__device__ __forceinline__ gdptr gmem(const double *p) { return (gdptr)p; }
__device__ __forceinline__ gdptr_w gmem_w(double *p) { return (gdptr_w)p; }
__device__ __forceinline__ gfptr gmem(const float *p) { return (gfptr)p; }

// Kernel-level entry points shared between translation units.
int launch_putin(x2g_grid *g, int isym, const double *fun, double *w1, double *w2);
int launch_putout(x2g_grid *g, double *fun, const double *work);
// putin of several functions (each on its own grid) in one launch
struct PadItem {
    const double *fun;
    double *w;
    int nni, nmu, isym;
};
int launch_putin_items(x2g_ctx *ctx, const std::vector<PadItem> &items);
// pinned host buffer of at least `bytes` for device -> host results
int pinned_out_reserve(x2g_ctx *ctx, size_t bytes);
// Look up the device address of the no-allocation scratch array on the
// context's device (current at the call); x2g_ctx_create calls it once.
int scratch_init(x2g_ctx *ctx);
// Sweep timing hooks (no-ops unless statistics are enabled on the context).
int stats_begin(x2g_ctx *ctx, hipEvent_t *a, hipEvent_t *b);
int stats_end(x2g_ctx *ctx, hipEvent_t a, hipEvent_t b, double bytes, int prec, int launches = 1);
int stats_flush(x2g_ctx *ctx);
}  // namespace x2g
