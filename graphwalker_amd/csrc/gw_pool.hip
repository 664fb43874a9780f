// gw_pool.hip -- walk pools held in host memory.
//
// The reference keeps its walk pool on disk: per-thread buffers of 4096 packed
// records per block spill to pool_<block>.walks files when full
// (WalkManager::writeWalks2Disk, walk.hpp:103-112) and are read back when the
// block runs (readWalksfromDisk, walk.hpp:139-157).  On one MI355X a walk
// takes 16 B of records plus up to 60 B of working copies and per-walk outputs
// in HBM, so pools beyond what the HBM budget leaves next to the graph stay in
// (pinned) host memory here: gw_walks_load_host registers the caller's
// records, and a run steps them window by window (gw_step.hip), copying the
// next window in on the copy stream and the previous window's per-walk outputs
// back (by walk id) on a stream of their own while one window steps.  Walk ids equal positions in the pool, so
// a window is one contiguous id range and trajectories equal those of the
// same records loaded into HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "gw_internal.h"

namespace {

int host_alloc(gw_ctx* ctx, void** p, uint64_t bytes, const char* what) {
    if (hipHostMalloc(p, std::max<uint64_t>(bytes, 16), hipHostMallocDefault) != hipSuccess) {
        *p = nullptr;
        return gw_fail(ctx, GW_ENOMEM, std::string("pinned host allocation failed: ") + what);
    }
    return GW_OK;
}

void host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

}  // namespace

void gw_pool_free(gw_ctx* ctx) {
    gw_pool* P = ctx->pool;
    if (!P) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    if (P->out_stream) (void)hipStreamSynchronize(P->out_stream);
    const uint64_t W = P->window;
    gw_free(ctx, P->d_in1, W * sizeof(gw_walk));
    gw_free(ctx, P->d_final1, W * sizeof(gw_walk));
    gw_free(ctx, P->d_hash1, W * 8);
    gw_free(ctx, P->d_visits1, W * 4);
    gw_free(ctx, P->d_events1, W * P->events1_cap * 8);
    host_free(P->h_final);
    host_free(P->h_hash);
    host_free(P->h_visits);
    host_free(P->h_events);
    host_free(P->h_cnt);
    for (auto* set : {&P->in_ev, &P->out_ev, &P->run_ev})
        for (auto& e : *set)
            if (e) (void)hipEventDestroy(e);
    if (P->out_stream) (void)hipStreamDestroy(P->out_stream);
    delete P;
    ctx->pool = nullptr;
}

int gw_pool_prepare(gw_ctx* ctx, uint32_t rec, uint32_t event_cap) {
    gw_pool& P = *ctx->pool;
    const uint64_t W = P.window;
    int rc;
    if (!ctx->copy_stream) GW_HIP(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    if (!P.out_stream) GW_HIP(ctx, hipStreamCreateWithFlags(&P.out_stream, hipStreamNonBlocking));
    for (auto* set : {&P.in_ev, &P.out_ev, &P.run_ev})
        for (auto& e : *set)
            if (!e) GW_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (!P.h_cnt && (rc = host_alloc(ctx, reinterpret_cast<void**>(&P.h_cnt), 2 * kCounters * 8, "pool counters")))
        return rc;
    if (!P.d_in1 && (rc = gw_alloc(ctx, reinterpret_cast<void**>(&P.d_in1), W * sizeof(gw_walk), "pool window")))
        return rc;
    if (!P.d_final1 && (rc = gw_alloc(ctx, reinterpret_cast<void**>(&P.d_final1), W * sizeof(gw_walk), "pool finals")))
        return rc;
    if (P.keep_final && !P.h_final &&
        (rc = host_alloc(ctx, reinterpret_cast<void**>(&P.h_final), P.n * sizeof(gw_walk), "pool final records")))
        return rc;
    if (rec & GW_REC_WALK_SUMMARY) {
        if (!P.d_hash1 && ((rc = gw_alloc(ctx, reinterpret_cast<void**>(&P.d_hash1), W * 8, "pool walk hash")) ||
                           (rc = gw_alloc(ctx, reinterpret_cast<void**>(&P.d_visits1), W * 4, "pool walk visits"))))
            return rc;
        if (!P.h_hash && ((rc = host_alloc(ctx, reinterpret_cast<void**>(&P.h_hash), P.n * 8, "pool walk hash")) ||
                          (rc = host_alloc(ctx, reinterpret_cast<void**>(&P.h_visits), P.n * 4, "pool walk visits"))))
            return rc;
    }
    if (rec & GW_REC_EVENTS) {
        if (P.events1_cap != event_cap) {
            gw_free(ctx, P.d_events1, W * P.events1_cap * 8);
            P.d_events1 = nullptr;
            P.events1_cap = 0;
            if ((rc = gw_alloc(ctx, reinterpret_cast<void**>(&P.d_events1), W * event_cap * 8, "pool event log")))
                return rc;
            P.events1_cap = event_cap;
        }
        if (P.h_events_cap != event_cap) {
            host_free(P.h_events);
            P.h_events = nullptr;
            P.h_events_cap = 0;
            if ((rc = host_alloc(ctx, reinterpret_cast<void**>(&P.h_events), P.n * event_cap * 8, "pool event log")))
                return rc;
            P.h_events_cap = event_cap;
        }
    }
    return GW_OK;
}

extern "C" {

int gw_walks_load_host(gw_ctx* ctx, const gw_walk* walks, uint64_t n, uint64_t window, uint32_t flags) {
    if (!ctx || !walks || n == 0 || window == 0 || (flags & ~GW_POOL_KEEP_FINAL))
        return gw_fail(ctx, GW_EINVAL, "gw_walks_load_host: bad arguments");
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_walks_load_host: no graph");
    if (n >= 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_walks_load_host: walk ids are 32-bit");
    uint32_t src_max = 0, v_max = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (walks[i].id != i || walks[i].vertex >= ctx->n)
            return gw_fail(ctx, GW_EINVAL, "gw_walks_load_host: walk " + std::to_string(i) +
                                               " has id != its position or a vertex out of range");
        src_max = std::max(src_max, walks[i].source);
        v_max = std::max(v_max, walks[i].vertex);
    }
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_pool_free(ctx);
    const uint64_t W = std::min(window, n);
    int rc = gw_reserve_walks(ctx, W);
    if (rc) return rc;
    gw_pool* P = new gw_pool();
    P->h_walks = walks;
    P->n = n;
    P->window = W;
    P->keep_final = (flags & GW_POOL_KEEP_FINAL) != 0;
    ctx->pool = P;
    ctx->walk_src_max = src_max;
    ctx->walk_v_max = v_max;
    ctx->field_walks.clear();
    ctx->field_vertex.clear();
    ctx->part_per_walk = 0;
    return GW_OK;
}

}  // extern "C"
