// Context lifetime, error strings and the scratch arena of the C ABI.
#include "nb_common.h"
#include "nb_ctx.h"
#include "nebula_hip.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

int32_t ctx_fail_code(nb_ctx* ctx, int32_t code, const char* msg)
{
    if (ctx) ctx->err = msg ? msg : "unknown error";
    return code;
}

int32_t ctx_fail(nb_ctx* ctx, const char* msg) { return ctx_fail_code(ctx, NB_ERR_ARG, msg); }

int ctx_reserve(nb_ctx* ctx, void** p, size_t* have, size_t bytes)
{
    if (bytes == 0) bytes = 1;
    if (*p && *have >= bytes) return 0;
    if (*p) {
        hipFree(*p);
        *p = nullptr;
        *have = 0;
        ctx->generation++;  // an address that a recording may hold is gone
    }
    size_t want = bytes + bytes / 8;  // headroom: sizes drift from step to step
    if (hipMalloc(p, want) != hipSuccess) {
        *p = nullptr;
        ctx_fail_code(ctx, NB_ERR_HIP, "hipMalloc of library scratch failed");
        return 1;
    }
    *have = want;
    return 0;
}

int ktime_begin(nb_ctx* ctx, hipStream_t s, const char* name)
{
    if (!ctx || !ctx->timing) return -1;
    if (ctx->ev_open.size() >= 512) ktime_flush(ctx);
    nb_timed t{name, nullptr, nullptr};
    for (hipEvent_t* e : {&t.start, &t.stop}) {
        if (!ctx->ev_pool.empty()) {
            *e = ctx->ev_pool.back();
            ctx->ev_pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            return -1;
        }
    }
    hipEventRecord(t.start, s);
    ctx->ev_open.push_back(t);
    return (int)ctx->ev_open.size() - 1;
}

void ktime_end(nb_ctx* ctx, hipStream_t s, int slot)
{
    if (!ctx || slot < 0 || slot >= (int)ctx->ev_open.size()) return;
    hipEventRecord(ctx->ev_open[slot].stop, s);
}

void ktime_flush(nb_ctx* ctx)
{
    for (nb_timed& t : ctx->ev_open) {
        float ms = 0.f;
        if (hipEventSynchronize(t.stop) == hipSuccess && hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            nb_kernel_time* k = nullptr;
            for (nb_kernel_time& x : ctx->ktimes)
                if (x.name == t.name) k = &x;
            if (!k) {
                ctx->ktimes.push_back(nb_kernel_time{t.name, 0, 0.0});
                k = &ctx->ktimes.back();
            }
            k->launches += 1;
            k->ms += ms;
        }
        ctx->ev_pool.push_back(t.start);
        ctx->ev_pool.push_back(t.stop);
    }
    ctx->ev_open.clear();
}

extern "C" int32_t nb_set_tile(nb_ctx* ctx, int32_t tile)
{
    if (!ctx || (tile != 64 && tile != 128 && tile != 256)) return ctx_fail(ctx, "nb_set_tile: 64, 128 or 256");
    ctx->tile = tile;
    return NB_OK;
}

extern "C" int32_t nb_timing(nb_ctx* ctx, int32_t enable)
{
    if (!ctx) return NB_ERR_ARG;
    ktime_flush(ctx);
    ctx->ktimes.clear();
    ctx->timing = enable ? 1 : 0;
    return NB_OK;
}

extern "C" int64_t nb_timing_report(nb_ctx* ctx, char* buf, int64_t cap)
{
    if (!ctx) return NB_ERR_ARG;
    ktime_flush(ctx);
    std::string out;
    char line[256];
    for (const nb_kernel_time& k : ctx->ktimes) {
        snprintf(line, sizeof(line), "%s %lld %.6f\n", k.name.c_str(), k.launches, k.ms);
        out += line;
    }
    if (buf && cap > 0) {
        const size_t n = std::min<size_t>((size_t)cap - 1, out.size());
        std::memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return (int64_t)out.size() + 1;
}

extern "C" nb_ctx* nb_ctx_create(int32_t device)
{
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    nb_ctx* ctx = new (std::nothrow) nb_ctx();
    if (!ctx) return nullptr;
    ctx->device = device;
    // the no-allocation entry point's device array; a failure only disables that entry point,
    // which then reports it
    if (static_scratch_init(device, &ctx->err)) ctx->err += " (nb_dt_minima_dev_static unavailable)";
    return ctx;
}

// every scratch buffer of the arena with its size, for release
#define NB_CTX_SCRATCH(X)                                                                                  \
    X(cell_id) X(keys_out) X(iota) X(order) X(cell_count) X(cell_start) X(gbuf) X(rowcnt) X(unsorted)     \
    X(pending) X(edbuf) X(ecount) X(eoff) X(efill) X(ebuf) X(tot) X(offs) X(tcount) X(dscratch) X(fscratch) \
    X(pack) X(part) X(posh) X(mlp_scratch) X(mlp_pack) X(mlp_part) X(segs) X(red) X(scal) X(cub_tmp) X(rm_gc) \
    X(rm_go) X(rm_ec) X(rm_eo) X(rm_to) X(ebuf_s) X(fx_tmp) X(fb_sz) X(fb_go) X(fb_to) X(segpart) X(xlist)

static void ctx_free_scratch(nb_ctx* ctx)
{
#define NB_FREE_ONE(F)          \
    if (ctx->F) hipFree(ctx->F); \
    ctx->F = nullptr;           \
    ctx->F##_bytes = 0;
    NB_CTX_SCRATCH(NB_FREE_ONE)
#undef NB_FREE_ONE
}

extern "C" void nb_ctx_destroy(nb_ctx* ctx)
{
    if (!ctx) return;
    ktime_flush(ctx);
    for (hipEvent_t e : ctx->ev_pool) hipEventDestroy(e);
    ctx_free_scratch(ctx);
    delete ctx;
}

extern "C" int32_t nb_ctx_trim(nb_ctx* ctx)
{
    if (!ctx) return NB_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return ctx_fail_code(ctx, NB_ERR_HIP, "nb_ctx_trim: device synchronisation failed");
    ctx_free_scratch(ctx);
    // what lived in the arena is gone: passes that read an earlier pass's scratch must be redone
    ctx->lists_ready = ctx->hydro_ready = ctx->hydro_mixed_ready = ctx->hydro64_ready = 0;
    ctx->fx_ready = ctx->fx_grid_from_order = 0;
    ctx->host_segs.clear();
    ctx->generation++;
    return NB_OK;
}

extern "C" const char* nb_last_error(const nb_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int32_t nb_abi_version(void) { return 4; }

extern "C" int64_t nb_ctx_generation(const nb_ctx* ctx) { return ctx ? (int64_t)ctx->generation : -1; }

extern "C" int64_t nb_ctx_scratch_bytes(const nb_ctx* ctx)
{
    if (!ctx) return 0;
    size_t s = 0;
#define NB_ADD_ONE(F) s += ctx->F##_bytes;
    NB_CTX_SCRATCH(NB_ADD_ONE)
#undef NB_ADD_ONE
    return (int64_t)s;
}
