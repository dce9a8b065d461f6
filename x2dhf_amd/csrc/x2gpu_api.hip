// Context, memory and grid management of the x2gpu C ABI.
#include "x2gpu_internal.h"
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <cstdlib>

namespace x2g {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int ws_reserve(x2g_ctx *ctx, int k, size_t bytes)
{
    if (ctx->ws[k].bytes >= bytes) return X2G_OK;
    // grow geometrically (a batch that grows one admission at a time then
    // pays few reallocations); the old buffer stays until the new one exists,
    // so a failed growth leaves the workspace as it was
    const size_t grown = ctx->ws[k].bytes + ctx->ws[k].bytes / 2;
    void *p = nullptr;
    size_t want = std::max(bytes, grown);
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        want = bytes;  // the exact size may still fit
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            set_error("workspace %d: hipMalloc(%zu) failed", k, bytes);
            return X2G_ENOMEM;
        }
    }
    // earlier launches may still read the old buffer
    if (ctx->ws[k].ptr) {
        X2G_HIP(hipStreamSynchronize(ctx->stream));
        X2G_HIP(hipFree(ctx->ws[k].ptr));
    }
    ctx->ws[k].ptr = p;
    ctx->ws[k].bytes = want;
    return X2G_OK;
}

int pinned_reserve(x2g_ctx *ctx, size_t bytes)
{
    if (ctx->pinned_bytes >= bytes) return X2G_OK;
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->pinned) X2G_HIP(hipHostFree(ctx->pinned));
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
    X2G_HIP(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
    ctx->pinned_bytes = bytes;
    return X2G_OK;
}

int pinned_out_reserve(x2g_ctx *ctx, size_t bytes)
{
    if (ctx->pinned_out_bytes >= bytes) return X2G_OK;
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->pinned_out) X2G_HIP(hipHostFree(ctx->pinned_out));
    ctx->pinned_out = nullptr;
    ctx->pinned_out_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 65536);
    X2G_HIP(hipHostMalloc(&ctx->pinned_out, want, hipHostMallocDefault));
    ctx->pinned_out_bytes = want;
    return X2G_OK;
}

int stage_bytes(x2g_ctx *ctx, const void *src, size_t bytes, const void **dev_out)
{
    if (ctx->cap_arena) {
        // graph capture: the descriptors belong to the graph's own arena
        const size_t off = (ctx->cap_off + 255) & ~size_t(255);
        if (off + bytes > ctx->cap_bytes) ctx->cap_overflow = true;
        X2G_CHECK(off + bytes <= ctx->cap_bytes, "launch descriptors exceed the graph arena (%zu bytes)",
                  ctx->cap_bytes);
        if (ctx->cap_host.size() < off + bytes) ctx->cap_host.resize(off + bytes);
        std::memcpy(ctx->cap_host.data() + off, src, bytes);
        *dev_out = ctx->cap_arena + off;
        ctx->cap_off = off + bytes;
        return X2G_OK;
    }
    const size_t cap = std::min(ctx->ws[2].bytes, ctx->pinned_bytes);
    X2G_CHECK(bytes <= cap, "launch descriptors (%zu bytes) exceed the staging ring", bytes);
    size_t off = (ctx->stage_off + 255) & ~size_t(255);
    if (off + bytes > cap) {
        // wrap: every copy and launch that used the ring has to be done
        const int rc = stage_flush(ctx);
        if (rc) return rc;
        X2G_HIP(hipStreamSynchronize(ctx->stream));
        off = 0;
    }
    std::memcpy((char *)ctx->pinned + off, src, bytes);
    if (ctx->stage_hi == ctx->stage_lo) ctx->stage_lo = off;
    ctx->stage_hi = off + bytes;
    *dev_out = (const char *)ctx->ws[2].ptr + off;
    ctx->stage_off = off + bytes;
    return X2G_OK;
}

int stage_flush(x2g_ctx *ctx)
{
    if (ctx->stage_hi > ctx->stage_lo && !ctx->cap_arena) {
        const size_t lo = ctx->stage_lo;
        X2G_HIP(hipMemcpyAsync((char *)ctx->ws[2].ptr + lo, (char *)ctx->pinned + lo, ctx->stage_hi - lo,
                               hipMemcpyHostToDevice, ctx->stream));
    }
    ctx->stage_lo = ctx->stage_hi = 0;
    return X2G_OK;
}

static hipEvent_t take_event(x2g_ctx *ctx)
{
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int stats_begin(x2g_ctx *ctx, hipEvent_t *a, hipEvent_t *b)
{
    *a = *b = nullptr;
    if (!ctx->stats_on) return X2G_OK;
    *a = take_event(ctx);
    *b = take_event(ctx);
    if (!*a || !*b) {
        set_error("hipEventCreate failed");
        return X2G_EHIP;
    }
    X2G_HIP(hipEventRecord(*a, ctx->stream));
    return X2G_OK;
}

int stats_end(x2g_ctx *ctx, hipEvent_t a, hipEvent_t b, double bytes, int prec, int launches)
{
    if (!ctx->stats_on || !a) return X2G_OK;
    X2G_HIP(hipEventRecord(b, ctx->stream));
    ctx->pending.push_back({a, b, bytes, prec, launches});
    if (ctx->pending.size() >= 4096) return stats_flush(ctx);
    return X2G_OK;
}

int stats_flush(x2g_ctx *ctx)
{
    if (ctx->pending.empty()) return X2G_OK;
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &p : ctx->pending) {
        float ms = 0.0f;
        X2G_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        ctx->st_launch[p.prec] += p.launches;
        ctx->st_ms[p.prec] += ms;
        ctx->st_bytes[p.prec] += p.bytes;
        ctx->ev_pool.push_back(p.a);
        ctx->ev_pool.push_back(p.b);
    }
    ctx->pending.clear();
    return X2G_OK;
}

}  // namespace x2g

using namespace x2g;

extern "C" int x2g_stats_enable(x2g_ctx *ctx, int on)
{
    X2G_CHECK(ctx, "x2g_stats_enable: null ctx");
    int rc = stats_flush(ctx);
    if (rc) return rc;
    ctx->stats_on = on != 0;
    for (int k = 0; k < 2; ++k) ctx->st_launch[k] = ctx->st_ms[k] = ctx->st_bytes[k] = 0;
    return X2G_OK;
}

extern "C" int x2g_stats_get(x2g_ctx *ctx, double *out)
{
    X2G_CHECK(ctx && out, "x2g_stats_get: null argument");
    int rc = stats_flush(ctx);
    if (rc) return rc;
    for (int k = 0; k < 2; ++k) {
        out[3 * k + 0] = ctx->st_launch[k];
        out[3 * k + 1] = ctx->st_ms[k];
        out[3 * k + 2] = ctx->st_bytes[k];
    }
    return X2G_OK;
}

static constexpr size_t kStageBytes = 16u << 20;  // launch-descriptor staging ring

extern "C" int x2g_version(void) { return 100; }

extern "C" const char *x2g_last_error(void) { return x2g::g_err; }

extern "C" int x2g_ctx_create(int device, x2g_ctx **out)
{
    X2G_CHECK(out, "x2g_ctx_create: null out");
    *out = nullptr;
    int ndev = 0;
    X2G_HIP(hipGetDeviceCount(&ndev));
    X2G_CHECK(device >= 0 && device < ndev, "x2g_ctx_create: device %d of %d", device, ndev);
    X2G_HIP(hipSetDevice(device));
    x2g_ctx *c = new x2g_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        set_error("hipGetDeviceProperties failed");
        return X2G_EHIP;
    }
    c->num_cu = prop.multiProcessorCount;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, device) != hipSuccess) optin = 0;
    size_t lds = std::max<size_t>((size_t)prop.sharedMemPerBlock, (size_t)std::max(optin, 0));
    // gfx950 provides 160 KB of LDS per compute unit
    c->max_lds = std::min<size_t>(lds, 160 * 1024);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreateWithFlags failed");
        return X2G_EHIP;
    }
    const char *gr = getenv("X2G_GRAPHS");
    c->graphs_on = gr && atoi(gr) != 0;
    int rc = ws_reserve(c, 2, kStageBytes);
    if (!rc) rc = pinned_reserve(c, kStageBytes);
    if (!rc) rc = scratch_init(c);
    if (rc) {
        x2g_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return X2G_OK;
}

void x2g::graphs_clear(x2g_ctx *ctx)
{
    for (auto &kv : ctx->graphs) {
        (void)hipGraphExecDestroy(kv.second.exec);
        (void)hipFree(kv.second.arena);
    }
    ctx->graphs.clear();
    ctx->graph_seen.clear();
}

extern "C" int x2g_ctx_destroy(x2g_ctx *ctx)
{
    if (!ctx) return X2G_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    graphs_clear(ctx);
    for (auto &p : ctx->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    for (auto &w : ctx->ws)
        if (w.ptr) (void)hipFree(w.ptr);
    for (auto &c : ctx->cache) (void)hipFree(c.second);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->pinned_out) (void)hipHostFree(ctx->pinned_out);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return X2G_OK;
}

extern "C" int x2g_ctx_sync(x2g_ctx *ctx)
{
    X2G_CHECK(ctx, "x2g_ctx_sync: null ctx");
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    return X2G_OK;
}

extern "C" void *x2g_ctx_stream(x2g_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int x2g_mem_info(x2g_ctx *ctx, size_t *free_bytes, size_t *total_bytes)
{
    X2G_CHECK(ctx && free_bytes && total_bytes, "x2g_mem_info: null argument");
    X2G_HIP(hipSetDevice(ctx->device));
    X2G_HIP(hipMemGetInfo(free_bytes, total_bytes));
    *free_bytes += ctx->cached_bytes;  // blocks this context keeps for reuse
    return X2G_OK;
}

// Freed blocks stay with the context for reuse (a job that ends and the next
// one that starts on the same grid exchange the same sizes), up to this many
// bytes; beyond it the largest cached blocks go back to the runtime.
static constexpr size_t kCacheLimit = (size_t)32 << 30;

static void cache_trim(x2g_ctx *ctx, size_t limit)
{
    if (ctx->cached_bytes <= limit) return;
    (void)hipStreamSynchronize(ctx->stream);
    while (ctx->cached_bytes > limit && !ctx->cache.empty()) {
        auto it = std::prev(ctx->cache.end());
        (void)hipFree(it->second);
        ctx->cached_bytes -= it->first;
        ctx->cache.erase(it);
    }
}

extern "C" int x2g_malloc(x2g_ctx *ctx, size_t bytes, void **dptr)
{
    X2G_CHECK(ctx && dptr, "x2g_malloc: null argument");
    *dptr = nullptr;
    if (bytes == 0) return X2G_OK;
    const size_t want = (bytes + 255) & ~size_t(255);
    // a cached block of the same size or up to 1/8 larger
    auto it = ctx->cache.lower_bound(want);
    if (it != ctx->cache.end() && it->first <= want + want / 8) {
        *dptr = it->second;
        ctx->live[*dptr] = it->first;
        ctx->cached_bytes -= it->first;
        ctx->cache.erase(it);
        return X2G_OK;
    }
    if (hipMalloc(dptr, want) != hipSuccess) {
        (void)hipGetLastError();
        // give the cache back and try once more
        cache_trim(ctx, 0);
        if (hipMalloc(dptr, want) != hipSuccess) {
            (void)hipGetLastError();
            *dptr = nullptr;
            set_error("hipMalloc(%zu) failed", bytes);
            return X2G_ENOMEM;
        }
    }
    ctx->live[*dptr] = want;
    return X2G_OK;
}

extern "C" int x2g_free(x2g_ctx *ctx, void *dptr)
{
    X2G_CHECK(ctx, "x2g_free: null ctx");
    if (!dptr) return X2G_OK;
    auto it = ctx->live.find(dptr);
    X2G_CHECK(it != ctx->live.end(), "x2g_free: %p was not allocated by x2g_malloc of this context", dptr);
    // kept for reuse: later work on the block is ordered after earlier work
    // by the context's stream, so no synchronisation is needed here
    ctx->cache.emplace(it->second, dptr);
    ctx->cached_bytes += it->second;
    ctx->live.erase(it);
    cache_trim(ctx, kCacheLimit);
    return X2G_OK;
}

extern "C" int x2g_h2d(x2g_ctx *ctx, void *dst_d, const void *src_h, size_t bytes)
{
    X2G_CHECK(ctx && ((dst_d && src_h) || bytes == 0), "x2g_h2d: null argument");
    if (!bytes) return X2G_OK;
    X2G_HIP(hipMemcpyAsync(dst_d, src_h, bytes, hipMemcpyHostToDevice, ctx->stream));
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    return X2G_OK;
}

extern "C" int x2g_d2h(x2g_ctx *ctx, void *dst_h, const void *src_d, size_t bytes)
{
    X2G_CHECK(ctx && ((dst_h && src_d) || bytes == 0), "x2g_d2h: null argument");
    if (!bytes) return X2G_OK;
    X2G_HIP(hipMemcpyAsync(dst_h, src_d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    X2G_HIP(hipStreamSynchronize(ctx->stream));
    return X2G_OK;
}

extern "C" int x2g_d2d(x2g_ctx *ctx, void *dst_d, const void *src_d, size_t bytes)
{
    X2G_CHECK(ctx && ((dst_d && src_d) || bytes == 0), "x2g_d2d: null argument");
    if (!bytes) return X2G_OK;
    X2G_HIP(hipMemcpyAsync(dst_d, src_d, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return X2G_OK;
}

extern "C" int x2g_grid_create(x2g_ctx *ctx, int nni, int nmu, const double *coef, const double *suppl,
                               const double *vec, x2g_grid **out)
{
    X2G_CHECK(ctx && coef && suppl && vec && out, "x2g_grid_create: null argument");
    X2G_CHECK(nni >= 7 && nmu >= 11, "x2g_grid_create: grid %dx%d too small", nni, nmu);
    X2G_CHECK((long)nni * nmu < (1L << 30), "x2g_grid_create: grid too large");
    *out = nullptr;
    static std::atomic<unsigned long long> next_uid{1};
    x2g_grid *g = new x2g_grid();
    g->ctx = ctx;
    g->uid = next_uid++;
    g->nni = nni;
    g->nmu = nmu;
    std::memcpy(g->coef, coef, sizeof(g->coef));
    const size_t n = (size_t)nni * nmu;
    int rc = X2G_OK;
    auto alloc = [&](void **p, size_t bytes) {
        if (rc) return;
        rc = x2g_malloc(ctx, bytes, p);
    };
    alloc((void **)&g->suppl, 14 * n * sizeof(double));
    alloc((void **)&g->bmu_orb, nmu * sizeof(double));
    alloc((void **)&g->bmu_pot, nmu * sizeof(double));
    alloc((void **)&g->dnu, nni * sizeof(double));
    alloc((void **)&g->bmu_orb_f, nmu * sizeof(float));
    alloc((void **)&g->bmu_pot_f, nmu * sizeof(float));
    alloc((void **)&g->dnu_f, nni * sizeof(float));
    alloc((void **)&g->vxi, nmu * sizeof(double));
    alloc((void **)&g->vxisq, nmu * sizeof(double));
    alloc((void **)&g->veta, nni * sizeof(double));
    alloc((void **)&g->vetasq, nni * sizeof(double));
    if (rc) {
        x2g_grid_destroy(g);
        return rc;
    }
    // b(mu) and d(nu) carry 1-D content (initSuppl, initCBAllocArrays.F90:1779-1781)
    std::vector<double> bo(nmu), bp(nmu), dn(nni);
    std::vector<float> bof(nmu), bpf(nmu), dnf(nni);
    for (int mu = 0; mu < nmu; ++mu) {
        bo[mu] = suppl[0 * n + (size_t)mu * nni];
        bp[mu] = suppl[1 * n + (size_t)mu * nni];
        bof[mu] = (float)bo[mu];
        bpf[mu] = (float)bp[mu];
    }
    for (int nu = 0; nu < nni; ++nu) {
        dn[nu] = suppl[2 * n + nu];
        dnf[nu] = (float)dn[nu];
    }
    if (!rc) rc = x2g_h2d(ctx, g->suppl, suppl, 14 * n * sizeof(double));
    if (!rc) rc = x2g_h2d(ctx, g->bmu_orb, bo.data(), nmu * sizeof(double));
    if (!rc) rc = x2g_h2d(ctx, g->bmu_pot, bp.data(), nmu * sizeof(double));
    if (!rc) rc = x2g_h2d(ctx, g->dnu, dn.data(), nni * sizeof(double));
    if (!rc) rc = x2g_h2d(ctx, g->bmu_orb_f, bof.data(), nmu * sizeof(float));
    if (!rc) rc = x2g_h2d(ctx, g->bmu_pot_f, bpf.data(), nmu * sizeof(float));
    if (!rc) rc = x2g_h2d(ctx, g->dnu_f, dnf.data(), nni * sizeof(float));
    if (!rc) rc = x2g_h2d(ctx, g->vxi, vec, nmu * sizeof(double));
    if (!rc) rc = x2g_h2d(ctx, g->vxisq, vec + nmu, nmu * sizeof(double));
    if (!rc) rc = x2g_h2d(ctx, g->veta, vec + 2 * nmu, nni * sizeof(double));
    if (!rc) rc = x2g_h2d(ctx, g->vetasq, vec + 2 * nmu + nni, nni * sizeof(double));
    if (rc) {
        x2g_grid_destroy(g);
        return rc;
    }
    *out = g;
    return X2G_OK;
}

extern "C" int x2g_grid_destroy(x2g_grid *g)
{
    if (!g) return X2G_OK;
    x2g_ctx *c = g->ctx;
    void *ptrs[] = {g->suppl, g->bmu_orb, g->bmu_pot, g->dnu, g->bmu_orb_f, g->bmu_pot_f,
                    g->dnu_f, g->vxi, g->vxisq, g->veta, g->vetasq};
    for (void *p : ptrs)
        if (p) x2g_free(c, p);
    for (auto &kv : g->levels) {
        if (kv.second.off) x2g_free(c, kv.second.off);
        if (kv.second.pts) x2g_free(c, kv.second.pts);
        if (kv.second.poff) x2g_free(c, kv.second.poff);
        if (kv.second.pops) x2g_free(c, kv.second.pops);
    }
    delete g;
    return X2G_OK;
}

extern "C" const double *x2g_grid_suppl(x2g_grid *g, int k)
{
    if (!g || k < 0 || k >= 14) return nullptr;
    return g->suppl + (size_t)k * g->n();
}
