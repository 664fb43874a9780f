// gw_step.hip -- step kernels: the device form of exec_updates + updateByWalk
// (src/engine/graphwalker.hpp:257-270, src/walks/randomwalkwith*.hpp).
//
// PLAIN       one walk per thread, each thread runs its walk to completion over
//             the whole resident graph.  The straightforward baseline: lanes of a
//             wavefront whose walks end early idle until the longest walk ends.
// PERSISTENT  a grid sized to the machine; each wavefront keeps 64 walks in
//             flight and refills a lane as soon as its walk ends, from a
//             wave-private chunk of walk indices (one global atomic per chunk).
//             Uneven, geometric walk lengths then cost no idle lanes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gw_device.h"
#include "gw_internal.h"

namespace gw {
namespace {

constexpr uint32_t kChunk = 512;  // walk indices a wavefront claims per atomic

// One step of a walk on the whole graph.  Returns true when the walk has
// ended (the final vertex/hop are left in v/hop, as the reference leaves them
// in the record it drops).
template <int KIND, class Rec>
__device__ __forceinline__ bool step_once(const Prog& p, const Graph& g, uint32_t id, uint32_t source,
                                          uint32_t& v, uint32_t& hop, Rec& rec) {
    constexpr bool capped = (KIND == GW_JUMP || KIND == GW_STOP || KIND == GW_RESTART);
    if (capped && hop >= p.L) return true;
    rec.visit(v, hop);
    const u32x4 r = philox4x32_10(id, hop, kTagWalk, p.stream, p.k0, p.k1);
    const uint64_t b = g.off[v];
    const uint64_t deg = g.off[v + 1] - b;
    if (deg > 0 && static_cast<uint64_t>(r.x) >= p.stop_thr) {
        const uint64_t x = static_cast<uint64_t>(r.y) | (static_cast<uint64_t>(r.z) << 32);
        v = g.col[b + mulhi64(x, deg)];
    } else {
        if (KIND == GW_JUMP) v = mulhi32(r.w, p.nvertices);
        else if (KIND == GW_STOP || KIND == GW_PROB) return true;
        else v = source;
    }
    ++hop;
    ++rec.steps;
    if (KIND == GW_RESTART_JOINT) return hop % p.L == p.L - 1;
    if (capped) return hop >= p.L;
    return false;
}

template <int KIND, bool VC, bool SUM>
__global__ __launch_bounds__(256) void step_plain_kernel(Prog p, Graph g, const gw_walk* __restrict__ walks,
                                                        gw_walk* __restrict__ final_walks, uint64_t nw,
                                                        uint32_t* __restrict__ counts, uint64_t* __restrict__ hash,
                                                        uint32_t* __restrict__ visits,
                                                        unsigned long long* __restrict__ counters) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    Recorder<VC, SUM> rec{counts, 0u, 0u, kHashInit};
    if (i < nw) {
        gw_walk w = walks[i];
        advance<KIND>(p, g, w.id, w.source, w.vertex, w.hop, 0u, g.n, rec);
        final_walks[w.id] = w;
        if (SUM) {
            hash[w.id] = rec.hash;
            visits[w.id] = rec.visits;
        }
    }
    const uint32_t st = wave_sum(rec.steps);
    const uint32_t vi = wave_sum(rec.visits);
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd(&counters[0], static_cast<unsigned long long>(st));
        atomicAdd(&counters[1], static_cast<unsigned long long>(vi));
    }
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <int KIND, bool VC, bool SUM>
__global__ __launch_bounds__(256) void step_persistent_kernel(Prog p, Graph g, const gw_walk* __restrict__ walks,
                                                             gw_walk* __restrict__ final_walks, uint64_t nw, uint32_t* __restrict__ counts,
                                                             uint64_t* __restrict__ hash,
                                                             uint32_t* __restrict__ visits,
                                                             unsigned long long* __restrict__ counters) {
    const uint32_t lane = lane_id();
    uint64_t chunk_next = 0, chunk_end = 0;  // wave-uniform
    bool exhausted = false;                  // wave-uniform
    bool active = false;
    uint64_t slot = 0;
    uint32_t id = 0, source = 0, v = 0, hop = 0;
    uint32_t tot_steps = 0, tot_visits = 0;
    Recorder<VC, SUM> rec{counts, 0u, 0u, kHashInit};

    for (;;) {
        // Refill idle lanes from the wave's chunk; claim a new chunk when dry.
        uint64_t idle = __ballot(!active);
        while (idle && !exhausted) {
            if (chunk_next == chunk_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&counters[2], static_cast<unsigned long long>(kChunk));
                base = __shfl(base, 0, 64);
                if (base >= nw) {
                    exhausted = true;
                    break;
                }
                chunk_next = base;
                chunk_end = std::min<uint64_t>(base + kChunk, nw);
            }
            const uint32_t want = __popcll(idle);
            const uint64_t avail = chunk_end - chunk_next;
            const uint32_t take = avail < want ? static_cast<uint32_t>(avail) : want;
            const uint32_t r = rank_in(idle);
            if (!active && r < take) {
                slot = chunk_next + r;
                const gw_walk w = walks[slot];
                id = w.id;
                source = w.source;
                v = w.vertex;
                hop = w.hop;
                active = true;
                rec.visits = 0;
                rec.steps = 0;
                rec.hash = kHashInit;
            }
            chunk_next += take;
            idle = __ballot(!active);
        }
        if (__ballot(active) == 0) break;
        if (active) {
            if (step_once<KIND>(p, g, id, source, v, hop, rec)) {
                final_walks[id] = gw_walk{id, source, v, hop};
                if (SUM) {
                    hash[id] = rec.hash;
                    visits[id] = rec.visits;
                }
                tot_steps += rec.steps;
                tot_visits += rec.visits;
                active = false;
            }
        }
    }
    const uint32_t st = wave_sum(tot_steps);
    const uint32_t vi = wave_sum(tot_visits);
    if (lane == 0) {
        atomicAdd(&counters[0], static_cast<unsigned long long>(st));
        atomicAdd(&counters[1], static_cast<unsigned long long>(vi));
    }
}

struct Launch {
    Prog p;
    Graph g;
    gw_ctx* ctx;
    uint32_t sched;
    uint32_t tb;
    uint32_t grid_persistent;
};

template <int KIND, bool VC, bool SUM>
hipError_t launch3(const Launch& L) {
    gw_ctx* c = L.ctx;
    if (L.sched == GW_SCHED_PLAIN) {
        const uint32_t grid = static_cast<uint32_t>((c->nwalks + L.tb - 1) / L.tb);
        step_plain_kernel<KIND, VC, SUM><<<dim3(grid), dim3(L.tb), 0, c->stream>>>(
            L.p, L.g, c->d_walks, c->d_final, c->nwalks, c->d_counts, c->d_hash, c->d_visits, c->d_counters);
    } else {
        step_persistent_kernel<KIND, VC, SUM><<<dim3(L.grid_persistent), dim3(L.tb), 0, c->stream>>>(
            L.p, L.g, c->d_walks, c->d_final, c->nwalks, c->d_counts, c->d_hash, c->d_visits, c->d_counters);
    }
    return hipGetLastError();
}

template <int KIND>
hipError_t launch2(const Launch& L, uint32_t rec) {
    const bool vc = rec & GW_REC_VERTEX_COUNT, sum = rec & GW_REC_WALK_SUMMARY;
    if (vc && sum) return launch3<KIND, true, true>(L);
    if (vc) return launch3<KIND, true, false>(L);
    if (sum) return launch3<KIND, false, true>(L);
    return launch3<KIND, false, false>(L);
}

hipError_t launch(const Launch& L, uint32_t kind, uint32_t rec) {
    switch (kind) {
    case GW_JUMP: return launch2<GW_JUMP>(L, rec);
    case GW_STOP: return launch2<GW_STOP>(L, rec);
    case GW_PROB: return launch2<GW_PROB>(L, rec);
    case GW_RESTART_JOINT: return launch2<GW_RESTART_JOINT>(L, rec);
    default: return launch2<GW_RESTART>(L, rec);
    }
}

}  // namespace
}  // namespace gw

int gw_step_run(gw_ctx* ctx, const gw_program* prog, gw_stats* stats) {
    using namespace gw;
    gw_stats st{};
    uint32_t sched = prog->schedule == GW_SCHED_AUTO ? GW_SCHED_PERSISTENT : prog->schedule;
    if (sched != GW_SCHED_PLAIN && sched != GW_SCHED_PERSISTENT)
        return gw_fail(ctx, GW_EINVAL, "gw_run: schedule not available");
    st.schedule = sched;
    if ((prog->recorder & GW_REC_WALK_SUMMARY) && !ctx->d_hash) {
        int rc;
        if ((rc = gw_alloc(ctx, (void**)&ctx->d_hash, ctx->walk_cap * 8, "walk hash"))) return rc;
        if ((rc = gw_alloc(ctx, (void**)&ctx->d_visits, ctx->walk_cap * 4, "walk visits"))) return rc;
    }
    Launch L;
    L.ctx = ctx;
    L.sched = sched;
    L.p = Prog{prog->L, prog->nvertices, prog->stop_thr, static_cast<uint32_t>(prog->seed),
               static_cast<uint32_t>(prog->seed >> 32), prog->stream};
    L.g = Graph{ctx->d_off, ctx->d_col, ctx->n};
    L.tb = prog->block_threads ? prog->block_threads : 256u;
    if (L.tb % 64 != 0 || L.tb > 256) return gw_fail(ctx, GW_EINVAL, "gw_run: block_threads must be a multiple of 64, at most 256");
    const uint32_t wps = prog->waves_per_simd ? prog->waves_per_simd : 8u;
    const uint64_t waves = static_cast<uint64_t>(std::max(1u, ctx->cu_count)) * 4u * wps;
    L.grid_persistent = static_cast<uint32_t>(std::max<uint64_t>(1, waves * 64 / L.tb));

    GW_HIP(ctx, hipMemsetAsync(ctx->d_counters, 0, 4 * sizeof(unsigned long long), ctx->stream));
    if (prog->recorder & GW_REC_VERTEX_COUNT) GW_HIP(ctx, hipMemsetAsync(ctx->d_counts, 0, ctx->n * 4ull, ctx->stream));
    GW_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if (ctx->nwalks) {
        hipError_t e = launch(L, prog->kind, prog->recorder);
        if (e != hipSuccess) return gw_hip_check(ctx, e, "step kernel launch");
        st.launches = 1;
    }
    GW_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    GW_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
    float ms = 0.f;
    GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    unsigned long long cnt[4];
    GW_HIP(ctx, hipMemcpy(cnt, ctx->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    st.steps = cnt[0];
    st.visits = cnt[1];
    st.rounds = 1;
    st.step_ms = ms;
    st.total_ms = ms;
    if (stats) *stats = st;
    return GW_OK;
}
