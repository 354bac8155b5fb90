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
    if (ctx) ctx->device = device;
    return ctx;
}

extern "C" void nb_ctx_destroy(nb_ctx* ctx)
{
    if (!ctx) return;
    ktime_flush(ctx);
    for (hipEvent_t e : ctx->ev_pool) hipEventDestroy(e);
    void* bufs[] = {ctx->cell_id, ctx->keys_out, ctx->iota, ctx->order, ctx->cell_count, ctx->cell_start,
                    ctx->gbuf, ctx->rowcnt, ctx->unsorted, ctx->pending, ctx->edbuf, ctx->ecount, ctx->eoff, ctx->efill, ctx->ebuf, ctx->tot,
                    ctx->offs, ctx->tcount, ctx->dscratch, ctx->fscratch, ctx->pack, ctx->part, ctx->posh, ctx->mlp_scratch, ctx->mlp_pack, ctx->mlp_part, ctx->segs, ctx->red, ctx->scal, ctx->cub_tmp,
                    ctx->rm_gc, ctx->rm_go, ctx->rm_ec, ctx->rm_eo, ctx->rm_to, ctx->ebuf_s, ctx->fx_tmp,
                    ctx->fb_sz, ctx->fb_go, ctx->fb_to, ctx->segpart, ctx->xlist};
    for (void* b : bufs)
        if (b) hipFree(b);
    delete ctx;
}

extern "C" const char* nb_last_error(const nb_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int32_t nb_abi_version(void) { return 4; }

extern "C" int64_t nb_ctx_scratch_bytes(const nb_ctx* ctx)
{
    if (!ctx) return 0;
    size_t s = ctx->cell_id_bytes + ctx->keys_out_bytes + ctx->iota_bytes + ctx->order_bytes +
               ctx->cell_count_bytes + ctx->cell_start_bytes + ctx->gbuf_bytes + ctx->rowcnt_bytes + ctx->unsorted_bytes + ctx->pending_bytes + ctx->edbuf_bytes +
               ctx->ecount_bytes + ctx->eoff_bytes + ctx->efill_bytes + ctx->ebuf_bytes + ctx->tot_bytes +
               ctx->offs_bytes + ctx->tcount_bytes + ctx->dscratch_bytes + ctx->fscratch_bytes + ctx->pack_bytes + ctx->part_bytes + ctx->posh_bytes + ctx->mlp_scratch_bytes + ctx->mlp_pack_bytes + ctx->mlp_part_bytes + ctx->segs_bytes +
               ctx->red_bytes + ctx->scal_bytes + ctx->cub_tmp_bytes + ctx->rm_gc_bytes + ctx->rm_go_bytes +
               ctx->rm_ec_bytes + ctx->rm_eo_bytes + ctx->rm_to_bytes + ctx->ebuf_s_bytes +
               ctx->fx_tmp_bytes + ctx->fb_sz_bytes + ctx->fb_go_bytes + ctx->fb_to_bytes +
               ctx->segpart_bytes + ctx->xlist_bytes;
    return (int64_t)s;
}
