// gw_step.hip -- step kernels: the device form of exec_updates + updateByWalk
// (src/engine/graphwalker.hpp:257-311, src/walks/randomwalkwith*.hpp).
//
// PLAIN       one walk per thread, run to completion over the whole resident
//             CSR.  The straightforward baseline: lanes whose walks end early
//             idle until the longest walk of their wavefront ends, and every
//             step makes two dependent random reads (offsets, then column).
// PERSISTENT  a grid sized to the machine; each wavefront keeps 64 walks in
//             flight and refills a lane as soon as its walk ends, from a
//             wave-private chunk of walk indices (one global atomic per chunk).
//             With the successor-header layout a step is one random read.
// BLOCKED     the reference's block loop on the device: walks are bucketed by
//             block (stable counting sort, gw_bucket.hip), each block's bucket
//             steps while its walks stay inside the block, walks that leave are
//             re-bucketed for the next round, finished walks drop out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "gw_device.h"
#include "gw_internal.h"

namespace gw {
namespace {

constexpr uint32_t kChunk = 512;           // default walk indices a wavefront claims per atomic
constexpr int kRecKeyLog = 1 << 16;        // internal recorder bit: source counts by sort-reduce
constexpr int kRecRegion = 1 << 17;        // internal recorder bit: visits straight into per-source regions
constexpr uint32_t kDead = 0xFFFFFFFFu;    // vertex of a finished walk in the blocked schedule

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <int KIND, int REC>
__global__ __launch_bounds__(256) void step_plain_kernel(Prog p, Graph g, const gw_walk* __restrict__ walks,
                                                        gw_walk* __restrict__ final_walks, uint64_t nw,
                                                        uint32_t* __restrict__ counts, uint64_t* __restrict__ hash,
                                                        uint32_t* __restrict__ visits,
                                                        unsigned long long* __restrict__ counters, EventLog log,
                                                        SourceTable table, KeyLog klog) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    constexpr bool SUM = (REC & GW_REC_WALK_SUMMARY) != 0;
    constexpr bool KL = (REC & kRecKeyLog) != 0;
    constexpr bool RG = (REC & kRecRegion) != 0;
    __shared__ uint64_t s_slab[4][2];
    if (KL && (threadIdx.x & 63u) == 0) {
        s_slab[threadIdx.x >> 6][0] = 0;
        s_slab[threadIdx.x >> 6][1] = 0;
    }
    __shared__ RegionBuf s_rbuf[RG ? 4 : 1];
    if (RG && (threadIdx.x & 63u) == 0) region_buf_init(&s_rbuf[threadIdx.x >> 6]);
    Recorder<(REC & GW_REC_VERTEX_COUNT) != 0, SUM, (REC & GW_REC_EVENTS) != 0, (REC & GW_REC_SOURCE_COUNT) != 0, KL, RG> rec(
        counts, log, table, klog, s_slab[threadIdx.x >> 6], RG ? &s_rbuf[threadIdx.x >> 6] : nullptr);
    if (i < nw) {
        gw_walk w = walks[i];
        const uint32_t o = w.id - p.id_base;
        rec.id = o;
        rec.source = w.source;
        if (RG) rec.field_v = w.source < klog.nf ? klog.field_vertex[w.source] : 0xFFFFFFFFu;
        advance<KIND>(p, g, w.id, w.source, w.vertex, w.hop, 0u, g.n, rec);
        if (final_walks) final_walks[o] = w;
        if (SUM) {
            hash[o] = rec.hash;
            visits[o] = rec.visits;
        }
    }
    if (KL) keylog_close(klog, s_slab[threadIdx.x >> 6], threadIdx.x & 63u);
    if (RG) region_close(klog, &s_rbuf[threadIdx.x >> 6]);
    const uint32_t st = wave_sum(rec.steps);
    const uint32_t vi = wave_sum(rec.visits);
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd(&counters[0], static_cast<unsigned long long>(st));
        atomicAdd(&counters[1], static_cast<unsigned long long>(vi));
    }
}

// GEN: the walks are those of gw_walks_init_sources, started from gen.tab
// (walk i: id i, hop 0, source entry i / gen.wps) instead of walk records.  A
// wavefront keeps the entries of two consecutive sources in scalar registers
// -- its walk indices [seg_lo, seg_hi), split at seg_mid -- so a walk start
// waits on no load: a walk-record read and the row read of its vertex would
// put two dependent memory latencies in front of the first step, and batched
// multi-source walks end every two or three steps.
template <int KIND, int REC, bool EXT, bool GEN = false>
__global__ __launch_bounds__(256) void step_persistent_kernel(Prog p, Graph g, const gw_walk* __restrict__ walks,
                                                             gw_walk* __restrict__ final_walks, uint64_t nw,
                                                             uint32_t* __restrict__ counts,
                                                             uint64_t* __restrict__ hash,
                                                             uint32_t* __restrict__ visits,
                                                             unsigned long long* __restrict__ counters, EventLog log,
                                                        SourceTable table, KeyLog klog, WalkGen gen = WalkGen{}) {
    const uint32_t lane = lane_id();
    uint64_t chunk_next = 0, chunk_end = 0;  // wave-uniform
    bool exhausted = false;                  // wave-uniform
    bool active = false;
    State s{};
    uint32_t seg_lo = 0, seg_mid = 0, seg_hi = 0;  // wave-uniform (GEN)
    SourceStart e0{}, e1{};                         // wave-uniform (GEN)
    uint32_t tot_steps = 0, tot_visits = 0;
    constexpr bool SUM = (REC & GW_REC_WALK_SUMMARY) != 0;
    constexpr bool KL = (REC & kRecKeyLog) != 0;
    constexpr bool RG = (REC & kRecRegion) != 0;
    __shared__ uint64_t s_slab[4][2];
    if (KL && (threadIdx.x & 63u) == 0) {
        s_slab[threadIdx.x >> 6][0] = 0;
        s_slab[threadIdx.x >> 6][1] = 0;
    }
    __shared__ RegionBuf s_rbuf[RG ? 4 : 1];
    if (RG && (threadIdx.x & 63u) == 0) region_buf_init(&s_rbuf[threadIdx.x >> 6]);
    Recorder<(REC & GW_REC_VERTEX_COUNT) != 0, SUM, (REC & GW_REC_EVENTS) != 0, (REC & GW_REC_SOURCE_COUNT) != 0, KL, RG, RG,
             RG && restarts<KIND>()> rec(
        counts, log, table, klog, s_slab[threadIdx.x >> 6], RG ? &s_rbuf[threadIdx.x >> 6] : nullptr);

    for (;;) {
        // Refill idle lanes from the wave's chunk; claim a new chunk when dry.
        uint64_t idle = __ballot(!active);
        while (idle && !exhausted) {
            if (chunk_next == chunk_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&counters[2], static_cast<unsigned long long>(p.chunk));
                base = lane_value(base, 0u);
                if (base >= nw) {
                    exhausted = true;
                    break;
                }
                chunk_next = base;
                chunk_end = std::min<uint64_t>(base + p.chunk, nw);
            }
            if (GEN && (chunk_next < seg_lo || chunk_next >= seg_hi)) {
                // the entries of the source holding chunk_next and the next one
                const uint32_t f = static_cast<uint32_t>(chunk_next) / gen.wps;
                seg_lo = f * gen.wps;
                seg_mid = seg_lo + gen.wps;
                seg_hi = f + 1 < gen.nsrc ? seg_mid + gen.wps : seg_mid;
                e0 = gen.tab[f];
                e1 = gen.tab[f + 1 < gen.nsrc ? f + 1 : f];
            }
            const uint32_t want = __popcll(idle);
            uint64_t avail = chunk_end - chunk_next;
            if (GEN) avail = std::min<uint64_t>(avail, seg_hi - chunk_next);
            const uint32_t take = avail < want ? static_cast<uint32_t>(avail) : want;
            const uint32_t r = rank_in(idle);
            if (GEN) {
                if (!active && r < take) {
                    const uint32_t i = static_cast<uint32_t>(chunk_next) + r;
                    const bool hi = i >= seg_mid;
                    s.id = i;
                    s.source = hi ? e1.field : e0.field;
                    s.v = hi ? e1.vertex : e0.vertex;
                    s.hop = 0;
                    if (EXT) {
                        s.row = hi ? e1.row : e0.row;
                        s.deg = hi ? e1.deg : e0.deg;
                        if (restarts<KIND>()) {
                            s.srow = hi ? e1.srow : e0.srow;
                            s.sdeg = hi ? e1.sdeg : e0.sdeg;
                        }
                    }
                    rec.id = i - p.id_base;
                    rec.source = s.source;
                    if (RG) rec.field_v = hi ? e1.fvertex : e0.fvertex;
                    start_walk(p, s);
                    active = true;
                    rec.visits = 0;
                    rec.steps = 0;
                    rec.hash = kHashInit;
                }
                chunk_next += take;
                idle = __ballot(!active);
                continue;
            }
            // A walk whose vertex (or restart target) is not in the graph is
            // skipped and counted: gw_run checks walks on the host, but the
            // walks of gw_run_noalloc are device data it never reads.
            constexpr bool kRestarts = KIND == GW_RESTART || KIND == GW_RESTART_JOINT;
            gw_walk w{};
            bool bad = false;
            if (!active && r < take) {
                w = walks[chunk_next + r];
                bad = w.vertex >= g.n || (kRestarts && w.source >= g.n);
                if (bad) atomicAdd(&counters[kCntBadWalks], 1ull);
            }
            if (!active && r < take && !bad) {
                s.id = w.id;
                rec.id = w.id - p.id_base;
                rec.source = w.source;
                if (RG) rec.field_v = w.source < klog.nf ? klog.field_vertex[w.source] : 0xFFFFFFFFu;
                s.source = w.source;
                s.v = w.vertex;
                s.hop = w.hop;
                load_row<EXT>(g, s);
                set_source_row<KIND, EXT>(g, s);
                start_walk(p, s);
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
            if (step_once<KIND, EXT>(p, g, s, rec)) {
                if (final_walks) final_walks[s.id - p.id_base] = gw_walk{s.id, s.source, s.v, s.hop};
                if (SUM) {
                    hash[s.id - p.id_base] = rec.hash;
                    visits[s.id - p.id_base] = rec.visits;
                }
                tot_steps += rec.steps;
                tot_visits += rec.visits;
                active = false;
            }
        }
        // the step's visit into its source's region, with every lane present
        if constexpr (RG) rec.commit();
    }
    if (KL) keylog_close(klog, s_slab[threadIdx.x >> 6], lane);
    if (RG) rec.close_regions();
    workgroup_add2(&counters[0], &counters[1], tot_steps, tot_visits);
}

constexpr uint32_t kMaxBlocks = 2048;

__device__ __forceinline__ uint32_t block_bounds(const uint32_t* bf, uint32_t nb, uint32_t v, uint32_t& lo,
                                                 uint32_t& hi) {
    uint32_t a = 0, b = nb;  // bf[a] <= v < bf[b]
    while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (bf[mid] <= v) a = mid; else b = mid;
    }
    lo = bf[a];
    hi = bf[a + 1];
    return a;
}

// A resident block of a host-resident graph, as the step code reads it:
// offsets and columns rebased so that off[v] and col[pos] take the graph's
// absolute vertex ids and edge positions (valid for v in the block and pos in
// its rows only).
struct BlockView {
    const uint64_t* off;
    const uint32_t* col;
};

// One block round over the buckets in [walks, walks + nw), back to back in
// block order.  STREAM: the graph is host-resident and each walk reads its
// block through views[block] (a device cache slot); the range holds only
// buckets of resident blocks.  Persistent wavefronts take walks in bucket order (so
// the blocks being stepped at any moment are one or two consecutive ones),
// and a walk steps while its vertex stays inside its block.  A walk that ends
// writes its final record and leaves a dead slot (vertex = kDead, dropped by
// the next bucketing); a walk that leaves the block is written back in place
// with its new vertex/hop for the next round.
template <int KIND, int REC, bool EXT, bool STREAM>
__global__ __launch_bounds__(256) void step_block_kernel(Prog p, Graph g, const uint32_t* __restrict__ block_first,
                                                        uint32_t nb, const BlockView* __restrict__ views,
                                                        gw_walk* __restrict__ walks, uint64_t nw,
                                                        gw_walk* __restrict__ final_walks,
                                                        uint32_t* __restrict__ counts, uint64_t* __restrict__ hash,
                                                        uint32_t* __restrict__ visits,
                                                        unsigned long long* __restrict__ counters, EventLog log,
                                                        SourceTable table, KeyLog klog) {
    __shared__ uint32_t s_bf[kMaxBlocks + 1];
    for (uint32_t i = threadIdx.x; i <= nb; i += blockDim.x) s_bf[i] = block_first[i];
    __syncthreads();
    const uint32_t lane = lane_id();
    uint64_t chunk_next = 0, chunk_end = 0;  // wave-uniform
    bool exhausted = false;                  // wave-uniform
    bool active = false;
    uint64_t slot = 0;
    uint32_t lo = 0, hi = 0, visits0 = 0;
    State s{};
    Graph gl = g;  // the graph as this lane's walk sees it (its block's view when streaming)
    uint32_t tot_steps = 0, tot_visits = 0;
    constexpr bool SUM = (REC & GW_REC_WALK_SUMMARY) != 0;
    constexpr bool KL = (REC & kRecKeyLog) != 0;
    constexpr bool RG = (REC & kRecRegion) != 0;
    __shared__ uint64_t s_slab[4][2];
    if (KL && (threadIdx.x & 63u) == 0) {
        s_slab[threadIdx.x >> 6][0] = 0;
        s_slab[threadIdx.x >> 6][1] = 0;
    }
    __shared__ RegionBuf s_rbuf[RG ? 4 : 1];
    if (RG && (threadIdx.x & 63u) == 0) region_buf_init(&s_rbuf[threadIdx.x >> 6]);
    Recorder<(REC & GW_REC_VERTEX_COUNT) != 0, SUM, (REC & GW_REC_EVENTS) != 0, (REC & GW_REC_SOURCE_COUNT) != 0, KL, RG> rec(
        counts, log, table, klog, s_slab[threadIdx.x >> 6], RG ? &s_rbuf[threadIdx.x >> 6] : nullptr);

    for (;;) {
        uint64_t idle = __ballot(!active);
        while (idle && !exhausted) {
            if (chunk_next == chunk_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&counters[2], static_cast<unsigned long long>(p.chunk));
                base = lane_value(base, 0u);
                if (base >= nw) {
                    exhausted = true;
                    break;
                }
                chunk_next = base;
                chunk_end = std::min<uint64_t>(base + p.chunk, nw);
            }
            const uint32_t want = __popcll(idle);
            const uint64_t avail = chunk_end - chunk_next;
            const uint32_t take = avail < want ? static_cast<uint32_t>(avail) : want;
            const uint32_t r = rank_in(idle);
            if (!active && r < take) {
                slot = chunk_next + r;
                const gw_walk w = walks[slot];
                s.id = w.id;
                rec.id = w.id - p.id_base;
                rec.source = w.source;
                if (RG) rec.field_v = w.source < klog.nf ? klog.field_vertex[w.source] : 0xFFFFFFFFu;
                s.source = w.source;
                s.v = w.vertex;
                s.hop = w.hop;
                const uint32_t bi = block_bounds(s_bf, nb, s.v, lo, hi);
                if (STREAM) {
                    const BlockView bv = views[bi];
                    gl.off = bv.off;
                    gl.col = bv.col;
                }
                load_row<EXT>(gl, s);
                set_source_row<KIND, EXT>(gl, s);
                start_walk(p, s);
                if (SUM) {
                    rec.hash = hash[s.id - p.id_base];
                    rec.visits = visits[s.id - p.id_base];
                } else {
                    rec.visits = 0;
                }
                visits0 = rec.visits;
                rec.steps = 0;
                active = true;
            }
            chunk_next += take;
            idle = __ballot(!active);
        }
        if (__ballot(active) == 0) break;
        if (active) {
            bool done = false, left = false;
            if (s.v - lo < hi - lo) {
                done = step_once<KIND, EXT>(p, gl, s, rec);
            } else {
                left = true;
            }
            if (done || left) {
                if (done && final_walks) final_walks[s.id - p.id_base] = gw_walk{s.id, s.source, s.v, s.hop};
                walks[slot] = gw_walk{s.id, s.source, done ? kDead : s.v, s.hop};
                if (SUM) {
                    hash[s.id - p.id_base] = rec.hash;
                    visits[s.id - p.id_base] = rec.visits;
                }
                tot_steps += rec.steps;
                tot_visits += rec.visits - visits0;
                active = false;
            }
        }
    }
    if (KL) keylog_close(klog, s_slab[threadIdx.x >> 6], lane);
    if (RG) region_close(klog, &s_rbuf[threadIdx.x >> 6]);
    const uint32_t st = wave_sum(tot_steps);
    const uint32_t vi = wave_sum(tot_visits);
    if (lane == 0) {
        atomicAdd(&counters[0], static_cast<unsigned long long>(st));
        atomicAdd(&counters[1], static_cast<unsigned long long>(vi));
    }
}

__global__ void source_start_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ field,
                                    uint32_t nsrc, const uint64_t* __restrict__ off, uint32_t n,
                                    const uint32_t* __restrict__ fvert, uint32_t nf, SourceStart* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsrc) return;
    const uint32_t v = src[k];
    const uint32_t f = field ? field[k] : v;
    const uint64_t b = off[v];
    SourceStart e{};
    e.vertex = v;
    e.field = f;
    e.fvertex = f < nf ? fvert[f] : 0xFFFFFFFFu;
    e.deg = static_cast<uint32_t>(off[v + 1] - b);
    e.row = b;
    if (f < n) {  // the field as a restart target
        e.srow = off[f];
        e.sdeg = static_cast<uint32_t>(off[f + 1] - e.srow);
    }
    out[k] = e;
}

__global__ void summary_init_kernel(const gw_walk* __restrict__ walks, uint64_t nw, uint32_t id_base,
                                    uint64_t* __restrict__ hash, uint32_t* __restrict__ visits) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= nw) return;
    const uint32_t id = walks[i].id - id_base;
    hash[id] = kHashInit;
    visits[id] = 0;
}

__global__ void build_ext_kernel(const uint64_t* __restrict__ off, const uint32_t* __restrict__ col, uint64_t m,
                                 Edge* __restrict__ ext, unsigned long long* __restrict__ overflow) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < m; e += stride) {
        const uint32_t d = col[e];
        const uint64_t b = off[d];
        const uint64_t deg = off[d + 1] - b;
        if (deg > 0xFFFFFFFFull) atomicAdd(overflow, 1ull);
        ext[e] = Edge{d, static_cast<uint32_t>(deg), b};
    }
}

struct Launch {
    Prog p;
    Graph g;
    gw_ctx* ctx;
    uint32_t sched;
    uint32_t tb;
    uint32_t grid_persistent;
    bool ext;
    EventLog log;
    SourceTable table;
    KeyLog klog;
    // blocked schedule: the buckets to step (back to back, block order)
    gw_walk* bucket;
    uint64_t bucket_n;
    const BlockView* views;  // non-null: host-resident graph, step through the block cache
    bool no_final;           // GW_REC_NO_FINAL: the kernels get no final-record array
    WalkGen gen;             // tab non-null: persistent walk starts from the source table
};

template <int KIND, int REC, bool EXT>
hipError_t launch4(const Launch& L) {
    gw_ctx* c = L.ctx;
    if (L.sched == GW_SCHED_PLAIN) {
        const uint32_t grid = static_cast<uint32_t>((c->nwalks + L.tb - 1) / L.tb);
        step_plain_kernel<KIND, REC><<<dim3(grid), dim3(L.tb), 0, c->stream>>>(
            L.p, L.g, c->d_walks, L.no_final ? nullptr : c->d_final, c->nwalks, c->d_counts, c->d_hash, c->d_visits, c->d_counters, L.log, L.table, L.klog);
    } else if (L.sched == GW_SCHED_PERSISTENT) {
        // generated walk starts: instantiated for the batched multi-source
        // recorder (direct regions) on the successor-header layout only
        if constexpr (REC == kRecRegion && EXT) {
            if (L.gen.tab) {
                step_persistent_kernel<KIND, REC, EXT, true><<<dim3(L.grid_persistent), dim3(L.tb), 0, c->stream>>>(
                    L.p, L.g, c->d_walks, L.no_final ? nullptr : c->d_final, c->nwalks, c->d_counts, c->d_hash,
                    c->d_visits, c->d_counters, L.log, L.table, L.klog, L.gen);
                return hipGetLastError();
            }
        }
        step_persistent_kernel<KIND, REC, EXT><<<dim3(L.grid_persistent), dim3(L.tb), 0, c->stream>>>(
            L.p, L.g, c->d_walks, L.no_final ? nullptr : c->d_final, c->nwalks, c->d_counts, c->d_hash, c->d_visits, c->d_counters, L.log, L.table, L.klog);
    } else if (L.views) {
        step_block_kernel<KIND, REC, false, true><<<dim3(L.grid_persistent), dim3(L.tb), 0, c->stream>>>(
            L.p, L.g, c->d_block_first, c->nblocks, L.views, L.bucket, L.bucket_n, L.no_final ? nullptr : c->d_final, c->d_counts,
            c->d_hash, c->d_visits, c->d_counters, L.log, L.table, L.klog);
    } else {
        step_block_kernel<KIND, REC, EXT, false><<<dim3(L.grid_persistent), dim3(L.tb), 0, c->stream>>>(
            L.p, L.g, c->d_block_first, c->nblocks, nullptr, L.bucket, L.bucket_n, L.no_final ? nullptr : c->d_final, c->d_counts,
            c->d_hash, c->d_visits, c->d_counters, L.log, L.table, L.klog);
    }
    return hipGetLastError();
}

template <int KIND, int REC>
hipError_t launch3(const Launch& L) {
    if (L.ext && L.sched != GW_SCHED_PLAIN) return launch4<KIND, REC, true>(L);
    return launch4<KIND, REC, false>(L);
}

// Recorder combinations that are instantiated; the event log always comes
// with the per-walk summary (its visit counts index the log).
template <int KIND>
hipError_t launch2(const Launch& L, uint32_t rec) {
    constexpr int VC = GW_REC_VERTEX_COUNT, SU = GW_REC_WALK_SUMMARY, EV = GW_REC_EVENTS, SC = GW_REC_SOURCE_COUNT;
    switch (rec) {
    case 0: return launch3<KIND, 0>(L);
    case VC: return launch3<KIND, VC>(L);
    case SU: return launch3<KIND, SU>(L);
    case VC | SU: return launch3<KIND, VC | SU>(L);
    case SU | EV: return launch3<KIND, SU | EV>(L);
    case VC | SU | EV: return launch3<KIND, VC | SU | EV>(L);
    case SC: return launch3<KIND, SC>(L);
    case SC | SU: return launch3<KIND, SC | SU>(L);
    case kRecKeyLog: return launch3<KIND, kRecKeyLog>(L);
    case kRecKeyLog | SU: return launch3<KIND, kRecKeyLog | SU>(L);
    case kRecRegion: return launch3<KIND, kRecRegion>(L);
    case kRecRegion | SU: return launch3<KIND, kRecRegion | SU>(L);
    default: return hipErrorInvalidValue;
    }
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

// Builds the successor-header edges once per graph.  Returns false (and
// leaves the CSR as the only layout) when it does not fit the HBM budget.
bool ensure_ext(gw_ctx* ctx) {
    if (ctx->d_ext) return true;
    if (ctx->ext_failed) return false;
    if (gw_alloc(ctx, (void**)&ctx->d_ext, std::max<uint64_t>(1, ctx->m) * sizeof(Edge), "successor-header edges")) {
        ctx->ext_failed = true;
        return false;
    }
    hipError_t e = hipMemsetAsync(ctx->d_counters + kCntExtOverflow, 0, 8, ctx->stream);
    if (e == hipSuccess && ctx->m) {
        build_ext_kernel<<<dim3(gw_grid(ctx->m, 256)), dim3(256), 0, ctx->stream>>>(
            ctx->d_off, ctx->d_col, ctx->m, reinterpret_cast<Edge*>(ctx->d_ext), ctx->d_counters + kCntExtOverflow);
        e = hipGetLastError();
    }
    unsigned long long overflow = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&overflow, ctx->d_counters + kCntExtOverflow, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess || overflow) {
        gw_free(ctx, ctx->d_ext, std::max<uint64_t>(1, ctx->m) * sizeof(Edge));
        ctx->d_ext = nullptr;
        ctx->ext_failed = true;
        return false;
    }
    return true;
}

int run_blocked(gw_ctx* ctx, Launch& L, const gw_program* prog, uint32_t rec, gw_stats& st) {
    // Working copy of the walks: the loaded records stay untouched so a run
    // can be repeated.
    gw_walk* a = ctx->d_walks_alt;
    gw_walk* b = ctx->d_work;
    GW_HIP(ctx, hipMemcpyAsync(a, ctx->d_walks, ctx->nwalks * sizeof(gw_walk), hipMemcpyDeviceToDevice, ctx->stream));
    if (rec & GW_REC_WALK_SUMMARY) {
        summary_init_kernel<<<dim3(static_cast<uint32_t>((ctx->nwalks + 255) / 256)), dim3(256), 0, ctx->stream>>>(
            ctx->d_walks, ctx->nwalks, L.p.id_base, ctx->d_hash, ctx->d_visits);
        GW_HIP(ctx, hipGetLastError());
    }
    const uint32_t nb = ctx->nblocks;
    uint64_t live = ctx->nwalks;
    std::vector<uint64_t> off;
    float ms = 0.f;
    double step_ms = 0, bucket_ms = 0;
    while (live > 0) {
        GW_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
        int rc = gw_bucket_sort(ctx, a, live, b, off);
        if (rc) return rc;
        GW_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
        GW_HIP(ctx, hipEventSynchronize(ctx->ev[3]));
        GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        bucket_ms += ms;
        live = off[nb];
        std::swap(a, b);  // buckets now in `a`
        if (live == 0) break;
        // One launch steps every bucket; wavefronts take walks in block order.
        GW_HIP(ctx, hipMemsetAsync(ctx->d_counters + kCntWorkCursor, 0, sizeof(unsigned long long), ctx->stream));
        GW_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
        L.bucket = a;
        L.bucket_n = live;
        hipError_t e = launch(L, prog->kind, rec);
        if (e != hipSuccess) return gw_hip_check(ctx, e, "block step launch");
        ++st.launches;
        GW_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
        GW_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
        GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        step_ms += ms;
        ++st.rounds;
    }
    st.step_ms = step_ms;
    st.bucket_ms = bucket_ms;
    return GW_OK;
}

// Host-resident graph: the blocked schedule over a device block cache.
// Each pass buckets the live walks by block.  While the resident blocks hold
// a worthwhile share of the walks (at least an eighth of the mean bucket of
// the blocks with walks), a drain pass steps just them, with no copies.
// Otherwise a load pass picks the blocks to step -- up to one per cache slot,
// by walk count, a resident block counting double (the device analogue of
// chooseBlock's max-walks choice, walk.hpp:204-252, with swapOut's
// fewest-walks eviction, graphwalker.hpp:236-251) -- copies the missing ones
// into slots of unchosen blocks on the copy stream, steps the buckets of the
// already-resident blocks while the copies run, then (after an event wait)
// the buckets of the newly loaded ones.  Block copies are the cost being
// minimised (PCIe moves ~57 GB/s, the step kernels much more), so each load
// is drained before the next.  Walks leave a block exactly as in the
// in-memory blocked schedule, so trajectories do not depend on the cache.
int run_streamed(gw_ctx* ctx, Launch& L, const gw_program* prog, uint32_t rec, gw_stats& st) {
    int rc = gw_stream_prepare(ctx);
    if (rc) return rc;
    gw_walk* a = ctx->d_walks_alt;
    gw_walk* b = ctx->d_work;
    GW_HIP(ctx, hipMemcpyAsync(a, ctx->d_walks, ctx->nwalks * sizeof(gw_walk), hipMemcpyDeviceToDevice, ctx->stream));
    if (rec & GW_REC_WALK_SUMMARY) {
        summary_init_kernel<<<dim3(static_cast<uint32_t>((ctx->nwalks + 255) / 256)), dim3(256), 0, ctx->stream>>>(
            ctx->d_walks, ctx->nwalks, L.p.id_base, ctx->d_hash, ctx->d_visits);
        GW_HIP(ctx, hipGetLastError());
    }
    const uint32_t nb = ctx->nblocks;
    const uint32_t nslots = static_cast<uint32_t>(ctx->slot_mem.size());
    BlockView* hv = static_cast<BlockView*>(ctx->h_views);
    L.views = static_cast<const BlockView*>(ctx->d_views);
    uint64_t live = ctx->nwalks;
    std::vector<uint64_t> off;
    std::vector<uint32_t> order, chosen, loaded;
    std::vector<char> in_round(nb, 0);
    const uint32_t policy = ctx->block_policy;
    std::vector<uint32_t> minhop;
    uint64_t coin_state = ctx->block_seed ^ 0x9E3779B97F4A7C15ull;
    auto coin = [&]() {  // splitmix64 -> uniform [0, 1)
        uint64_t z = (coin_state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return static_cast<double>((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
    };
    float ms = 0.f;
    double step_ms = 0, bucket_ms = 0, copy_ms = 0;
    // Steps the buckets of `blocks` (sorted), one launch per run of adjacent blocks.
    auto step_blocks = [&](const std::vector<uint32_t>& blocks) -> int {
        for (size_t i = 0; i < blocks.size();) {
            size_t j = i + 1;
            while (j < blocks.size() && blocks[j] == blocks[j - 1] + 1) ++j;
            const uint64_t first = off[blocks[i]], last = off[blocks[j - 1] + 1];
            i = j;
            if (last == first) continue;
            GW_HIP(ctx, hipMemsetAsync(ctx->d_counters + kCntWorkCursor, 0, sizeof(unsigned long long), ctx->stream));
            L.bucket = a + first;
            L.bucket_n = last - first;
            hipError_t e = launch(L, prog->kind, rec);
            if (e != hipSuccess) return gw_hip_check(ctx, e, "streamed step launch");
            ++st.launches;
        }
        return GW_OK;
    };
    while (live > 0) {
        GW_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
        if ((rc = gw_bucket_sort(ctx, a, live, b, off, policy != GW_BLOCKS_MAX_WALKS ? &minhop : nullptr))) return rc;
        GW_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
        GW_HIP(ctx, hipEventSynchronize(ctx->ev[3]));
        GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        bucket_ms += ms;
        live = off[nb];
        std::swap(a, b);  // buckets now in `a`
        if (live == 0) break;
        // drain pass: the resident blocks still hold enough walks
        uint64_t resident_walks = 0, with_walks = 0;
        chosen.clear();
        for (uint32_t k = 0; k < nb; ++k) {
            const uint64_t c = off[k + 1] - off[k];
            with_walks += c > 0;
            if (c > 0 && ctx->block_slot[k] >= 0) {
                resident_walks += c;
                chosen.push_back(k);
            }
        }
        if (!chosen.empty() && resident_walks * 8 * with_walks >= live) {
            GW_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
            if ((rc = step_blocks(chosen))) return rc;
            GW_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
            GW_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
            GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            step_ms += ms;
            ++st.rounds;
            continue;
        }
        // load pass: choose the blocks
        order.clear();
        for (uint32_t k = 0; k < nb; ++k)
            if (off[k + 1] > off[k]) order.push_back(k);
        // walks per block (MAX_WEIGHT: per hop of the block's least-advanced
        // walk, a min hop of 0 counting as 1), resident blocks counting double
        auto score = [&](uint32_t k) {
            const double w = static_cast<double>(off[k + 1] - off[k]) * (ctx->block_slot[k] >= 0 ? 2.0 : 1.0);
            return policy == GW_BLOCKS_MAX_WEIGHT ? w / std::max(1u, minhop[k]) : w;
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return score(x) > score(y); });
        if (policy == GW_BLOCKS_MIN_STEP && coin() < ctx->block_prob) {
            // the block holding the least-advanced walk goes first (the lowest
            // block id among equals, as blockWithMinStep scans)
            uint32_t best = order[0];
            for (uint32_t k : order)
                if (minhop[k] < minhop[best] || (minhop[k] == minhop[best] && k < best)) best = k;
            order.erase(std::find(order.begin(), order.end(), best));
            order.insert(order.begin(), best);
        }
        if (order.size() > nslots) order.resize(nslots);
        chosen.clear();
        loaded.clear();
        for (uint32_t k : order) in_round[k] = 1;
        // slots whose block is not in this round, emptiest first
        std::vector<uint32_t> spare;
        for (uint32_t sl = 0; sl < nslots; ++sl)
            if (ctx->slot_block[sl] < 0 || !in_round[ctx->slot_block[sl]]) spare.push_back(sl);
        size_t next_spare = 0;
        bool copies = false;
        for (uint32_t k : order) {
            if (ctx->block_slot[k] >= 0) {
                chosen.push_back(k);
                continue;
            }
            const uint32_t sl = spare[next_spare++];  // one slot per chosen block, so one is spare
            if (ctx->slot_block[sl] >= 0) ctx->block_slot[ctx->slot_block[sl]] = -1;
            ctx->slot_block[sl] = static_cast<int32_t>(k);
            ctx->block_slot[k] = static_cast<int32_t>(sl);
            const uint32_t lo = ctx->block_first[k], hi = ctx->block_first[k + 1];
            const uint64_t e0 = ctx->h_off[lo], e1 = ctx->h_off[hi];
            char* mem = static_cast<char*>(ctx->slot_mem[sl]);
            const uint64_t col_at = ((hi - lo + 1ull) * 8 + 255) / 256 * 256;  // columns after the offsets
            if (!copies) GW_HIP(ctx, hipEventRecord(ctx->copy_ev[1], ctx->copy_stream));
            copies = true;
            GW_HIP(ctx, hipMemcpyAsync(mem, ctx->h_off + lo, (hi - lo + 1ull) * 8, hipMemcpyHostToDevice,
                                       ctx->copy_stream));
            if (e1 > e0)
                GW_HIP(ctx, hipMemcpyAsync(mem + col_at, ctx->h_col + e0, (e1 - e0) * 4, hipMemcpyHostToDevice,
                                           ctx->copy_stream));
            st.h2d_bytes += (hi - lo + 1ull) * 8 + (e1 - e0) * 4;
            ++st.blocks_loaded;
            // rebased so the kernels index with absolute vertex ids / edge positions
            hv[k].off = reinterpret_cast<const uint64_t*>(reinterpret_cast<uintptr_t>(mem) - uintptr_t{lo} * 8);
            hv[k].col = reinterpret_cast<const uint32_t*>(reinterpret_cast<uintptr_t>(mem + col_at) -
                                                          static_cast<uintptr_t>(e0) * 4);
            loaded.push_back(k);
        }
        for (uint32_t k : order) in_round[k] = 0;
        if (copies) {
            GW_HIP(ctx, hipEventRecord(ctx->copy_ev[2], ctx->copy_stream));
            GW_HIP(ctx, hipEventRecord(ctx->copy_ev[0], ctx->copy_stream));
        }
        GW_HIP(ctx, hipMemcpyAsync(ctx->d_views, ctx->h_views, nb * sizeof(BlockView), hipMemcpyHostToDevice,
                                   ctx->stream));
        std::sort(chosen.begin(), chosen.end());
        std::sort(loaded.begin(), loaded.end());
        GW_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
        if ((rc = step_blocks(chosen))) return rc;  // overlaps the copies
        if (copies) GW_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->copy_ev[0], 0));
        if ((rc = step_blocks(loaded))) return rc;
        GW_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
        GW_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
        GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        step_ms += ms;
        if (copies) {
            GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->copy_ev[1], ctx->copy_ev[2]));
            copy_ms += ms;
        }
        ++st.rounds;
    }
    st.step_ms = step_ms;
    st.bucket_ms = bucket_ms;
    st.copy_ms = copy_ms;
    st.cache_slots = nslots;
    L.views = nullptr;
    return GW_OK;
}

// Key log for sort-reduce source counts: capacity for every visit the
// program can make (capped kinds: L per walk; uncapped: 64 per walk, more is
// an overflow error) plus slab padding.
uint32_t keylog_per_walk(const gw_program* prog) { return prog->kind == GW_PROB ? 64u : prog->L; }

int prepare_keylog(gw_ctx* ctx, const gw_program* prog, uint32_t sched, uint32_t grid_waves, uint32_t* slab) {
    const uint64_t per_walk = keylog_per_walk(prog);
    // Slab per wavefront: large for the persistent grid (few, long-lived
    // wavefronts), small for the plain grid (one short-lived wavefront per 64
    // walks).  Each launch leaves at most one partly used slab per wavefront.
    uint64_t launches = 1, waves = grid_waves;
    if (sched == GW_SCHED_PLAIN) {
        *slab = 64;
        waves = (ctx->nwalks + 63) / 64;
    } else if (sched == GW_SCHED_BLOCKED) {
        *slab = 256;
        launches = per_walk + 1;  // every live walk steps at least once per round
    } else {
        *slab = 1024;
    }
    const uint64_t cap = std::max<uint64_t>(1, ctx->nwalks) * per_walk + launches * waves * *slab + *slab;
    if (cap > ctx->keys_cap) {
        gw_free(ctx, ctx->d_keys, ctx->keys_cap * 8);
        ctx->d_keys = nullptr;
        ctx->keys_cap = 0;
        int rc = gw_alloc(ctx, (void**)&ctx->d_keys, cap * 8, "source key log");
        if (rc) return rc;
        ctx->keys_cap = cap;
    }
    const uint32_t key_bits = 64 - ctx->table_cbits;
    ctx->keys_sentinel = 1ull << key_bits;
    return GW_OK;
}

// Sizes, clears and records the layout of the (source, vertex) count table.
int prepare_source_table(gw_ctx* ctx, const gw_program* prog) {
    auto bits = [](uint64_t x) {
        uint32_t b = 0;
        while (b < 64 && (x >> b) != 0) ++b;
        return b;
    };
    const uint32_t vbits = std::max(1u, bits(ctx->n ? ctx->n - 1ull : 0ull));
    const uint32_t sbits = bits(ctx->walk_src_max);
    if (vbits + sbits > 63)
        return gw_fail(ctx, GW_EINVAL, "gw_run: source field and vertex ids need more than 63 key bits");
    uint32_t lg = prog->table_log2;
    // a host pool's windows add to one table: size it for the whole pool
    const uint64_t nw_all = ctx->pool ? ctx->pool->n : ctx->nwalks;
    if (lg == 0) lg = std::max(12u, std::min(31u, bits(std::max<uint64_t>(1, nw_all) * 4 - 1)));
    if (lg < 4 || lg > 34) return gw_fail(ctx, GW_EINVAL, "gw_run: table_log2 out of range [4, 34]");
    ctx->table_log2 = lg;
    ctx->table_vbits = vbits;
    ctx->table_cbits = 64 - vbits - sbits;
    if (prog->source_method != GW_SRC_HASH && !ctx->host_graph && !ctx->pool) return GW_OK;  // key log, no hash table
    // Keys of at most 44 bits pack with a 20-bit count into one word; longer
    // keys (a restart program's vertex-id source fields on a large graph)
    // take 16-byte entries {key + 1, count}, both words in one sector.
    const bool wide = vbits + sbits > 44;
    const uint64_t slots = 1ull << lg;
    const uint64_t bytes = slots * (wide ? 16 : 8);
    if (bytes != ctx->table_bytes) {
        gw_free(ctx, ctx->d_table, ctx->table_bytes);
        ctx->d_table = nullptr;
        ctx->table_slots = ctx->table_bytes = 0;
        int rc = gw_alloc(ctx, (void**)&ctx->d_table, bytes, "source-count table");
        if (rc) return rc;
        ctx->table_bytes = bytes;
    }
    ctx->table_slots = slots;
    ctx->table_wide = wide;
    return GW_OK;  // cleared by each run (a fill kernel, or a memset in the blocked schedule)
}

}  // namespace
}  // namespace gw

bool gw_ensure_ext(gw_ctx* ctx) { return gw::ensure_ext(ctx); }

namespace gw {
namespace {

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

}  // namespace

// Everything gw_run decides before it enqueues work.
struct RunPlan {
    gw_program prog{};
    uint32_t sched = 0, rec = 0;
    bool ext = false, part = false;
    bool direct = false;  // partitioned counts written by the step kernel itself (no key log, no partition pass)
    Launch L{};
    gw_stats st0{};  // schedule and layout
};

}  // namespace gw

struct gw_run_graph {
    gw::RunPlan plan;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> fp;             // what the graph's nodes point at and launch with
    unsigned long long* h_cnt = nullptr;  // pinned: the counters the graph copies back
};

void gw_run_graph_free(gw_ctx* ctx) {
    gw_run_graph* g = ctx->run_graph;
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->h_cnt) (void)hipHostFree(g->h_cnt);
    delete g;
    ctx->run_graph = nullptr;
}

namespace gw {
namespace {

// Validation and allocation: every buffer the run touches exists afterwards.
int plan_run(gw_ctx* ctx, const gw_program* prog, RunPlan& rp) {
    rp = RunPlan{};
    rp.prog = *prog;
    gw_stats& st = rp.st0;
    const uint32_t sched = prog->schedule != GW_SCHED_AUTO ? prog->schedule
                           : ctx->host_graph            ? GW_SCHED_BLOCKED
                                                        : GW_SCHED_PERSISTENT;
    if (sched != GW_SCHED_PLAIN && sched != GW_SCHED_PERSISTENT && sched != GW_SCHED_BLOCKED)
        return gw_fail(ctx, GW_EINVAL, "gw_run: unknown schedule");
    if (prog->layout > GW_LAYOUT_EXT) return gw_fail(ctx, GW_EINVAL, "gw_run: unknown layout");
    if (ctx->pool && (prog->recorder & GW_REC_SOURCE_COUNT) &&
        (prog->source_method == GW_SRC_GLOBAL || prog->source_method == GW_SRC_PARTITIONED))
        return gw_fail(ctx, GW_EINVAL, "gw_run: a host walk pool counts sources with GW_SRC_HASH");
    if (ctx->host_graph) {
        if (sched != GW_SCHED_BLOCKED)
            return gw_fail(ctx, GW_EINVAL, "gw_run: a host-resident graph runs the blocked schedule");
        if (prog->layout == GW_LAYOUT_EXT)
            return gw_fail(ctx, GW_EINVAL, "gw_run: a host-resident graph is read in the CSR layout");
        if ((prog->recorder & GW_REC_SOURCE_COUNT) &&
            (prog->source_method == GW_SRC_GLOBAL || prog->source_method == GW_SRC_PARTITIONED))
            return gw_fail(ctx, GW_EINVAL, "gw_run: a host-resident graph counts sources with GW_SRC_HASH");
    }
    if (sched == GW_SCHED_PLAIN && prog->layout == GW_LAYOUT_EXT)
        return gw_fail(ctx, GW_EINVAL, "gw_run: the plain schedule reads the CSR layout");
    st.schedule = sched;
    bool ext = false;
    if (sched != GW_SCHED_PLAIN && prog->layout != GW_LAYOUT_CSR && !ctx->host_graph) {
        ext = ensure_ext(ctx);
        if (!ext && prog->layout == GW_LAYOUT_EXT)
            return gw_fail(ctx, GW_ENOMEM, "gw_run: successor-header layout does not fit (HBM budget or degree >= 2^32)");
    }
    st.layout = ext ? GW_LAYOUT_EXT : GW_LAYOUT_CSR;
    uint32_t rec = prog->recorder & ~(GW_REC_NO_FINAL | GW_REC_EVENTS_LAST);
    if (rec & GW_REC_EVENTS) rec |= GW_REC_WALK_SUMMARY;
    if ((rec & GW_REC_EVENTS) && ctx->events_cap != prog->event_cap) {
        gw_free(ctx, ctx->d_events, ctx->walk_cap * ctx->events_cap * 8);
        ctx->d_events = nullptr;
        ctx->events_cap = 0;
        int rc = gw_alloc(ctx, (void**)&ctx->d_events, std::max<uint64_t>(1, ctx->walk_cap) * prog->event_cap * 8, "event log");
        if (rc) return rc;
        ctx->events_cap = prog->event_cap;
    }
    ctx->src_method = GW_SRC_HASH;
    uint32_t keylog_slab = 1024;
    if (rec & GW_REC_SOURCE_COUNT) {
        int rc = prepare_source_table(ctx, prog);
        if (rc) return rc;
        if (prog->source_method != GW_SRC_HASH && !ctx->host_graph && !ctx->pool) {
            if (prog->source_method == GW_SRC_PARTITIONED && ctx->field_walks.empty())
                return gw_fail(ctx, GW_EINVAL, "gw_run: GW_SRC_PARTITIONED needs source fields below 2^20");
            const uint32_t wps = prog->waves_per_simd ? prog->waves_per_simd : 8u;
            const uint32_t tb = prog->block_threads ? prog->block_threads : 256u;
            const uint32_t grid_waves = std::max(1u, ctx->cu_count) * 4u * wps / (tb / 64) * (tb / 64);
            if ((rc = prepare_keylog(ctx, prog, sched, grid_waves, &keylog_slab))) return rc;
            rec = (rec & ~GW_REC_SOURCE_COUNT) | kRecKeyLog;
            ctx->src_method = GW_SRC_GLOBAL;
        }
    }
    if ((rec & GW_REC_WALK_SUMMARY) && !ctx->d_hash) {
        int rc;
        if ((rc = gw_alloc(ctx, (void**)&ctx->d_hash, ctx->walk_cap * 8, "walk hash"))) return rc;
        if ((rc = gw_alloc(ctx, (void**)&ctx->d_visits, ctx->walk_cap * 4, "walk visits"))) return rc;
    }
    if (sched == GW_SCHED_BLOCKED && !ctx->d_work && ctx->walk_cap) {
        int rc = gw_alloc(ctx, (void**)&ctx->d_work, ctx->walk_cap * sizeof(gw_walk), "blocked work buffer");
        if (rc) return rc;
    }
    Launch& L = rp.L;
    L.ctx = ctx;
    L.sched = sched;
    L.ext = ext;
    const uint32_t chunk = prog->walk_chunk ? prog->walk_chunk : kChunk;
    if (chunk % 64 != 0 || chunk > 8192)
        return gw_fail(ctx, GW_EINVAL, "gw_run: walk_chunk must be a multiple of 64, at most 8192");
    L.p = Prog{prog->L, prog->nvertices, prog->stop_thr, static_cast<uint32_t>(prog->seed),
               static_cast<uint32_t>(prog->seed >> 32), prog->stream, 0u, chunk};
    L.g = Graph{ctx->d_off, ctx->d_col, reinterpret_cast<const Edge*>(ctx->d_ext), ctx->n};
    L.log = EventLog{ctx->d_events, (rec & GW_REC_EVENTS) ? prog->event_cap : 0u,
                     (rec & GW_REC_EVENTS) && (prog->recorder & GW_REC_EVENTS_LAST) ? 1u : 0u};
    L.table = SourceTable{ctx->d_table, ctx->d_counters + kCntTableOverflow, ctx->table_slots - 1, 64u - ctx->table_log2,
                          ctx->table_vbits, ctx->table_cbits, ctx->table_wide ? 1u : 0u};
    L.klog = KeyLog{};  // region fields stay null unless the run fills regions directly
    L.klog.keys = ctx->d_keys;
    L.klog.cursor = ctx->d_counters + kCntLogCursor;
    L.klog.overflow = ctx->d_counters + kCntLogOverflow;
    L.klog.cap = ctx->keys_cap;
    L.klog.sentinel = ctx->keys_sentinel;
    L.klog.vbits = ctx->table_vbits;
    L.klog.slab = keylog_slab;
    L.no_final = (prog->recorder & GW_REC_NO_FINAL) != 0;
    L.tb = prog->block_threads ? prog->block_threads : 256u;
    if (L.tb % 64 != 0 || L.tb > 256) return gw_fail(ctx, GW_EINVAL, "gw_run: block_threads must be a multiple of 64, at most 256");
    const uint32_t wps = prog->waves_per_simd ? prog->waves_per_simd : 8u;
    if (wps > 32) return gw_fail(ctx, GW_EINVAL, "gw_run: waves_per_simd must be at most 32");
    const uint64_t waves = static_cast<uint64_t>(std::max(1u, ctx->cu_count)) * 4u * wps;
    L.grid_persistent = static_cast<uint32_t>(std::max<uint64_t>(1, waves * 64 / L.tb));
    // Partitioned source counts: regions and pair buffers are sized before
    // the run so partition + aggregation follow the step kernel on the
    // stream with no host round trip in between.
    // The automatic method takes the partitioned aggregation while the mean
    // source region (walks x keys per walk) fits what the LDS tables absorb
    // without flushing most sources; heavier sources sort-reduce globally.
    bool part_ok = !ctx->field_walks.empty();
    if (part_ok && prog->source_method == GW_SRC_SORT) {
        uint64_t nonempty = 0;
        for (uint64_t w : ctx->field_walks) nonempty += w != 0;
        part_ok = ctx->nwalks * keylog_per_walk(prog) <= kPartAutoRegionKeys * std::max<uint64_t>(1, nonempty);
    }
    if ((rec & kRecKeyLog) && prog->source_method != GW_SRC_GLOBAL && part_ok) {
        int rc = gw_partlog_prepare(ctx, keylog_per_walk(prog), &rp.part);
        if (rc) return rc;
        if (!rp.part && prog->source_method == GW_SRC_PARTITIONED)
            return gw_fail(ctx, GW_ENOMEM, "gw_run: the partitioned key log does not fit the HBM budget");
        // A capped walk makes at most L visits, so regions sized walks x L
        // cannot overflow: the step kernel appends each visit to its source's
        // region directly.  Uncapped (GW_PROB) walks keep the key log, whose
        // global sort-reduce takes over when a region overflows.
        if (rp.part && prog->kind != GW_PROB) {
            rec = (rec & ~kRecKeyLog) | kRecRegion;
            rp.direct = true;
            L.klog.region_off = ctx->d_part_off;
            L.klog.region_end = ctx->d_part_off + 1;
            L.klog.region_n = ctx->d_part_cursor;
            L.klog.region_out = ctx->d_part_keys;
            L.klog.region_overflow = ctx->d_counters + kCntPartOverflow;
            L.klog.nf = ctx->part_fields;
            L.klog.field_vertex = ctx->d_part_fvert;
            L.klog.field_count = ctx->d_part_fcount;
        }
    }
    // walks of gw_walks_init_sources start from the per-source table
    L.gen = WalkGen{};
    if (sched == GW_SCHED_PERSISTENT && ext && rec == kRecRegion && ctx->gen_wps && !ctx->pool && !ctx->host_graph)
        L.gen = WalkGen{static_cast<const SourceStart*>(ctx->d_gen_table), ctx->gen_wps, ctx->gen_nsrc};
    rp.sched = sched;
    rp.rec = rec;
    return GW_OK;
}

// The per-source start entries of generated walks, rebuilt on the engine
// stream when the walks, the graph or the source fields changed since the
// last build (outside any captured graph: a replay calls it before launching).
int gen_table_ready(gw_ctx* ctx, const RunPlan& rp) {
    if (!rp.L.gen.tab) return GW_OK;
    const uint64_t built[3] = {ctx->graph_epoch, reinterpret_cast<uintptr_t>(rp.L.klog.field_vertex), rp.L.klog.nf};
    if (std::equal(built, built + 3, ctx->gen_built)) return GW_OK;
    source_start_kernel<<<dim3((ctx->gen_nsrc + 255) / 256), dim3(256), 0, ctx->stream>>>(
        ctx->d_gen_src, ctx->d_gen_field, ctx->gen_nsrc, ctx->d_off, ctx->n, rp.L.klog.field_vertex, rp.L.klog.nf,
        const_cast<SourceStart*>(rp.L.gen.tab));
    GW_HIP(ctx, hipGetLastError());
    std::copy(built, built + 3, ctx->gen_built);
    return GW_OK;
}

// The work of one run on the engine stream, for the persistent and plain
// schedules: counter reset, recorder clears (fill kernels, so a captured
// graph holds no memset nodes; skipped for the later windows of a host walk
// pool, which add to the first's), the step kernel, and the partitioned
// aggregation.  timed: HIP events around the step kernel and the aggregation.
int enqueue_run(gw_ctx* ctx, RunPlan& rp, bool timed, bool clear = true) {
    int rc = gw_partlog_reset(ctx, rp.part);
    if (rc) return rc;
    if (clear && (rp.rec & GW_REC_VERTEX_COUNT)) {
        fill_u32_kernel<<<dim3(gw_grid(ctx->n, 256)), dim3(256), 0, ctx->stream>>>(ctx->d_counts, ctx->n);
        GW_HIP(ctx, hipGetLastError());
    }
    if (clear && (rp.rec & GW_REC_SOURCE_COUNT)) {  // hash table
        const uint64_t words = ctx->table_bytes / 8;
        fill_u64_kernel<<<dim3(gw_grid(words, 256)), dim3(256), 0, ctx->stream>>>(ctx->d_table, words);
        GW_HIP(ctx, hipGetLastError());
    }
    if (timed) GW_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    if (ctx->nwalks) {
        hipError_t e = launch(rp.L, rp.prog.kind, rp.rec);
        if (e != hipSuccess) return gw_hip_check(ctx, e, "step kernel launch");
    }
    if (timed) GW_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    if (rp.part && (rc = gw_partlog_enqueue(ctx, !rp.direct))) return rc;
    if (timed) GW_HIP(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
    return GW_OK;
}

// Counters read back -> stats; source-count aggregation that needs the host.
int finish_run(gw_ctx* ctx, RunPlan& rp, const unsigned long long* cnt, gw_stats& st, double reduce_ms) {
    st.steps = cnt[kCntSteps];
    st.visits = cnt[kCntVisits];
    const uint32_t rec = rp.rec;
    if ((rec & GW_REC_SOURCE_COUNT) && cnt[kCntTableOverflow])
        return gw_fail(ctx, GW_ENOMEM, "gw_run: source-count table full (" + std::to_string(cnt[kCntTableOverflow]) +
                                           " visits not recorded); raise table_log2");
    if (rec & (kRecKeyLog | kRecRegion)) {
        if (cnt[kCntLogOverflow])
            return gw_fail(ctx, GW_ENOMEM, "gw_run: source key log full (walks longer than the log allows)");
        double rms = 0;
        bool done = false;
        int rc;
        if (rp.part) {
            rms = reduce_ms;
            double extra = 0;
            if ((rc = gw_partlog_finish(ctx, cnt, &done, &extra))) return rc;
            rms += extra;
            if (!done && rp.direct)  // regions hold every visit a capped walk can make
                return gw_fail(ctx, GW_ESTATE, "gw_run: a source's visits outgrew its region");
            if (!done && rp.prog.source_method == GW_SRC_PARTITIONED)
                return gw_fail(ctx, GW_ENOMEM, "gw_run: a source's visits outgrew its key-log region (GW_SRC_PARTITIONED)");
        }
        if (done) {
            ctx->src_method = GW_SRC_PARTITIONED;
            st.flushed_sources = static_cast<uint32_t>(ctx->part_flushed);
        } else {
            double gms = 0;
            if ((rc = gw_keylog_reduce(ctx, cnt[kCntLogCursor], &gms))) return rc;
            ctx->src_method = GW_SRC_GLOBAL;
            rms += gms;
        }
        st.total_ms += rms;
        st.reduce_ms = rms;
    }
    st.source_method = (rec & (GW_REC_SOURCE_COUNT | kRecKeyLog | kRecRegion)) ? ctx->src_method : 0u;
    return GW_OK;
}

// Every device pointer, size and launch parameter a captured run bakes into
// its graph: a replay is valid while the freshly planned run has the same.
std::vector<uint64_t> fingerprint(const gw_ctx* ctx, const RunPlan& rp) {
    auto p = [](const void* x) { return static_cast<uint64_t>(reinterpret_cast<uintptr_t>(x)); };
    return {p(ctx->d_walks), ctx->nwalks, p(ctx->d_final), p(ctx->d_hash), p(ctx->d_visits), p(ctx->d_counts),
            ctx->n, p(ctx->d_off), p(ctx->d_col), p(ctx->d_ext), p(ctx->d_events), ctx->events_cap,
            p(ctx->d_table), ctx->table_bytes, ctx->table_slots, ctx->table_log2, ctx->table_vbits, ctx->table_cbits,
            ctx->table_wide, p(ctx->d_keys), ctx->keys_cap, ctx->keys_sentinel, p(ctx->d_part_keys),
            p(ctx->d_part_counts), p(ctx->d_part_off), p(ctx->d_part_cursor), p(ctx->d_part_pairs), p(ctx->d_part_fvert),
            p(ctx->d_part_fcount), p(ctx->d_part_items), ctx->part_fields,
            ctx->part_total, p(ctx->d_split_counts), ctx->split_cap, p(ctx->d_counters), rp.part, rp.rec, rp.sched,
            rp.ext, rp.L.tb, rp.L.grid_persistent, rp.L.klog.slab, ctx->cu_count, rp.L.no_final,
            p(rp.L.gen.tab), rp.L.gen.wps, rp.L.gen.nsrc};
}

// Host walk pool: the run steps the pool window by window through two device
// slots.  While the engine stream steps window k in slot k & 1, the copy
// stream copies window k + 1's records into the other slot and the pool's
// output stream copies window k - 1's per-walk outputs out of it (the two
// directions of PCIe at once).  Window k + 2 waits for both.  Per-vertex counts and
// the source-count hash add up over the windows (cleared once); the device
// counters are reset per window and summed on the host one window behind, so
// the host never holds the engine stream between windows.
int run_pool(gw_ctx* ctx, const gw_program* prog, gw_stats* stats) {
    gw_pool& P = *ctx->pool;
    const uint64_t W = P.window, n = P.n;
    const uint64_t nwin = (n + W - 1) / W;
    P.final_valid = P.summary_valid = P.events_valid = false;
    ctx->nwalks = W;
    RunPlan rp;
    int rc = plan_run(ctx, prog, rp);
    if (rc) return rc;
    const uint32_t rec = rp.rec;
    const bool sum = (rec & GW_REC_WALK_SUMMARY) != 0;
    const uint64_t ecap = (rec & GW_REC_EVENTS) ? prog->event_cap : 0u;
    if ((rc = gw_pool_prepare(ctx, rec, static_cast<uint32_t>(ecap)))) return rc;
    struct Slot {
        gw_walk* in;
        gw_walk* fin;
        uint64_t* hash;
        uint32_t* vis;
        uint64_t* ev;
    };
    const Slot slot[2] = {{ctx->d_walks, ctx->d_final, ctx->d_hash, ctx->d_visits, ctx->d_events},
                          {P.d_in1, P.d_final1, P.d_hash1, P.d_visits1, P.d_events1}};
    struct Restore {  // the context's own arrays again on every exit
        gw_ctx* c;
        Slot s;
        uint64_t nw;
        ~Restore() {
            c->d_walks = s.in;
            c->d_final = s.fin;
            c->d_hash = s.hash;
            c->d_visits = s.vis;
            c->d_events = s.ev;
            c->nwalks = nw;
        }
    } restore{ctx, slot[0], W};
    hipStream_t cs = ctx->copy_stream, os = P.out_stream;
    gw_stats st = rp.st0;
    uint64_t bytes = 0;
    auto count_of = [&](uint64_t k) { return std::min(W, n - k * W); };
    auto copy_in = [&](uint64_t k) -> int {
        const uint64_t c = count_of(k);
        GW_HIP(ctx, hipMemcpyAsync(slot[k & 1].in, P.h_walks + k * W, c * sizeof(gw_walk), hipMemcpyHostToDevice, cs));
        GW_HIP(ctx, hipEventRecord(P.in_ev[k & 1], cs));
        bytes += c * sizeof(gw_walk);
        return GW_OK;
    };
    auto absorb = [&](uint64_t k) -> int {
        GW_HIP(ctx, hipEventSynchronize(P.run_ev[k & 1]));
        const unsigned long long* cnt = P.h_cnt + (k & 1) * kCounters;
        if ((rec & GW_REC_SOURCE_COUNT) && cnt[kCntTableOverflow])
            return gw_fail(ctx, GW_ENOMEM, "gw_run: source-count table full (" + std::to_string(cnt[kCntTableOverflow]) +
                                               " visits not recorded); raise table_log2");
        st.steps += cnt[kCntSteps];
        st.visits += cnt[kCntVisits];
        return GW_OK;
    };
    const bool blocked = rp.sched == GW_SCHED_BLOCKED;
    if (blocked) {  // the recorders are cleared once for all windows
        if (rec & GW_REC_VERTEX_COUNT) GW_HIP(ctx, hipMemsetAsync(ctx->d_counts, 0, ctx->n * 4ull, ctx->stream));
        if (rec & GW_REC_SOURCE_COUNT) GW_HIP(ctx, hipMemsetAsync(ctx->d_table, 0, ctx->table_bytes, ctx->stream));
    }
    GW_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    GW_HIP(ctx, hipStreamWaitEvent(cs, ctx->ev[4], 0));  // slot 0's arrays are free from here on
    const bool outputs = P.keep_final || sum || ecap;
    if ((rc = copy_in(0))) return rc;
    if (nwin > 1 && (rc = copy_in(1))) return rc;
    st.rounds = st.launches = 0;
    for (uint64_t k = 0; k < nwin; ++k) {
        const uint32_t s = static_cast<uint32_t>(k & 1);
        const uint64_t c = count_of(k), base = k * W;
        GW_HIP(ctx, hipStreamWaitEvent(ctx->stream, P.in_ev[s], 0));
        if (k >= 2 && outputs) GW_HIP(ctx, hipStreamWaitEvent(ctx->stream, P.out_ev[s], 0));
        ctx->d_walks = slot[s].in;
        ctx->d_final = slot[s].fin;
        ctx->d_hash = slot[s].hash;
        ctx->d_visits = slot[s].vis;
        ctx->d_events = slot[s].ev;
        ctx->nwalks = c;
        rp.L.p.id_base = static_cast<uint32_t>(base);
        rp.L.log.ev = slot[s].ev;
        if (blocked) {
            if ((rc = gw_partlog_reset(ctx, false))) return rc;
            gw_stats w = rp.st0;
            rc = ctx->host_graph ? run_streamed(ctx, rp.L, prog, rec, w) : run_blocked(ctx, rp.L, prog, rec, w);
            if (rc) return rc;
            st.rounds += w.rounds;
            st.launches += w.launches;
            st.step_ms += w.step_ms;
            st.bucket_ms += w.bucket_ms;
            st.h2d_bytes += w.h2d_bytes;
            st.blocks_loaded += w.blocks_loaded;
            st.copy_ms += w.copy_ms;
            st.cache_slots = w.cache_slots;
        } else {
            if ((rc = enqueue_run(ctx, rp, false, k == 0))) return rc;
            st.launches += 1;
            st.rounds += 1;
        }
        GW_HIP(ctx, hipMemcpyAsync(P.h_cnt + s * kCounters, ctx->d_counters, kCounters * 8, hipMemcpyDeviceToHost,
                                   ctx->stream));
        GW_HIP(ctx, hipEventRecord(P.run_ev[s], ctx->stream));
        // window k's per-walk outputs out on the output stream; window k + 2
        // into this slot on the copy stream
        if (outputs) {
            GW_HIP(ctx, hipStreamWaitEvent(os, P.run_ev[s], 0));
            if (P.keep_final) {
                GW_HIP(ctx, hipMemcpyAsync(P.h_final + base, slot[s].fin, c * sizeof(gw_walk), hipMemcpyDeviceToHost, os));
                bytes += c * sizeof(gw_walk);
            }
            if (sum) {
                GW_HIP(ctx, hipMemcpyAsync(P.h_hash + base, slot[s].hash, c * 8, hipMemcpyDeviceToHost, os));
                GW_HIP(ctx, hipMemcpyAsync(P.h_visits + base, slot[s].vis, c * 4, hipMemcpyDeviceToHost, os));
                bytes += c * 12;
            }
            if (ecap) {
                GW_HIP(ctx, hipMemcpyAsync(P.h_events + base * ecap, slot[s].ev, c * ecap * 8, hipMemcpyDeviceToHost, os));
                bytes += c * ecap * 8;
            }
            GW_HIP(ctx, hipEventRecord(P.out_ev[s], os));
        }
        if (k + 2 < nwin) {
            GW_HIP(ctx, hipStreamWaitEvent(cs, P.run_ev[s], 0));
            if ((rc = copy_in(k + 2))) return rc;
        }
        if (k >= 1 && (rc = absorb(k - 1))) return rc;
    }
    if ((rc = absorb(nwin - 1))) return rc;
    // the run ends when the last outputs are on the host
    if (outputs) GW_HIP(ctx, hipStreamWaitEvent(ctx->stream, P.out_ev[(nwin - 1) & 1], 0));
    GW_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    GW_HIP(ctx, hipStreamSynchronize(cs));
    GW_HIP(ctx, hipEventSynchronize(ctx->ev[5]));
    float ms = 0.f;
    GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
    st.total_ms = ms;
    if (!blocked) st.step_ms = ms;
    st.pool_windows = nwin;
    st.pool_bytes = bytes;
    st.source_method = (rec & GW_REC_SOURCE_COUNT) ? GW_SRC_HASH : 0u;
    ctx->src_method = GW_SRC_HASH;
    P.final_valid = P.keep_final && !rp.L.no_final;
    P.summary_valid = sum;
    P.events_valid = ecap != 0;
    if (stats) *stats = st;
    return GW_OK;
}

}  // namespace
}  // namespace gw

int gw_step_run(gw_ctx* ctx, const gw_program* prog, gw_stats* stats) {
    using namespace gw;
    if (ctx->pool) return run_pool(ctx, prog, stats);
    RunPlan rp;
    int rc = plan_run(ctx, prog, rp);
    if (rc) return rc;
    gw_stats st = rp.st0;
    float ms = 0.f;
    double reduce_ms = 0;
    if (rp.sched == GW_SCHED_BLOCKED) {
        if ((rc = gw_partlog_reset(ctx, rp.part))) return rc;
        if (rp.rec & GW_REC_VERTEX_COUNT) GW_HIP(ctx, hipMemsetAsync(ctx->d_counts, 0, ctx->n * 4ull, ctx->stream));
        if (rp.rec & GW_REC_SOURCE_COUNT) GW_HIP(ctx, hipMemsetAsync(ctx->d_table, 0, ctx->table_bytes, ctx->stream));
        GW_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
        rc = !ctx->nwalks     ? GW_OK
             : ctx->host_graph ? run_streamed(ctx, rp.L, prog, rp.rec, st)
                               : run_blocked(ctx, rp.L, prog, rp.rec, st);
        if (rc) return rc;
        GW_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
        if (rp.part && (rc = gw_partlog_enqueue(ctx, !rp.direct))) return rc;
        GW_HIP(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
    } else {
        if ((rc = gen_table_ready(ctx, rp)) || (rc = enqueue_run(ctx, rp, true))) return rc;
        st.launches = ctx->nwalks ? 1 : 0;
        st.rounds = 1;
    }
    // through pinned memory: a copy to pageable memory is staged by the
    // runtime before it returns
    if (!ctx->h_run_cnt &&
        hipHostMalloc(reinterpret_cast<void**>(&ctx->h_run_cnt), kCounters * 8, hipHostMallocDefault) != hipSuccess) {
        ctx->h_run_cnt = nullptr;
        return gw_fail(ctx, GW_ENOMEM, "gw_run: pinned counter buffer");
    }
    const unsigned long long* cnt = ctx->h_run_cnt;
    GW_HIP(ctx, hipMemcpyAsync(ctx->h_run_cnt, ctx->d_counters, kCounters * 8, hipMemcpyDeviceToHost, ctx->stream));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
    st.total_ms = ms;
    if (rp.sched != GW_SCHED_BLOCKED) st.step_ms = ms;
    if (rp.part) {
        GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[5], ctx->ev[6]));
        reduce_ms = ms;
    }
    if ((rc = finish_run(ctx, rp, cnt, st, reduce_ms))) return rc;
    ctx->finals_valid = !rp.L.no_final;
    if (stats) *stats = st;
    return GW_OK;
}

extern "C" {

int gw_run_capture(gw_ctx* ctx, const gw_program* prog) {
    using namespace gw;
    if (!ctx || !prog) return gw_fail(ctx, GW_EINVAL, "gw_run_capture: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_run_graph_free(ctx);
    if (ctx->pool) return gw_fail(ctx, GW_EINVAL, "gw_run_capture: a host walk pool runs window by window");
    // one ordinary run first: allocations settle and the run's own checks pass
    gw_stats st0;
    int rc = gw_run(ctx, prog, &st0);
    if (rc) return rc;
    RunPlan rp;
    if ((rc = plan_run(ctx, prog, rp))) return rc;
    if (rp.sched == GW_SCHED_BLOCKED)
        return gw_fail(ctx, GW_EINVAL, "gw_run_capture: the blocked schedule decides rounds on the host");
    if ((rp.rec & kRecKeyLog) && !rp.part)
        return gw_fail(ctx, GW_EINVAL, "gw_run_capture: the global sort-reduce is not captured (use partitioned or hash counts)");
    gw_run_graph* g = new gw_run_graph();
    if (hipHostMalloc(reinterpret_cast<void**>(&g->h_cnt), kCounters * 8, hipHostMallocDefault) != hipSuccess) {
        delete g;
        return gw_fail(ctx, GW_ENOMEM, "gw_run_capture: pinned counter buffer");
    }
    ctx->run_graph = g;
    g->plan = rp;
    if ((rc = gen_table_ready(ctx, rp))) return rc;
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    GW_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    rc = enqueue_run(ctx, g->plan, false);
    hipError_t e = rc ? hipSuccess : hipMemcpyAsync(g->h_cnt, ctx->d_counters, kCounters * 8, hipMemcpyDeviceToHost, ctx->stream);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
    g->graph = graph;
    if (rc || e != hipSuccess || e2 != hipSuccess) {
        gw_run_graph_free(ctx);
        if (rc) return rc;
        return gw_hip_check(ctx, e != hipSuccess ? e : e2, "gw_run_capture: stream capture");
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        gw_run_graph_free(ctx);
        return gw_hip_check(ctx, e, "gw_run_capture: graph instantiate");
    }
    g->fp = fingerprint(ctx, g->plan);
    return GW_OK;
}

int gw_run_replay(gw_ctx* ctx, gw_stats* stats) {
    using namespace gw;
    if (!ctx) return GW_EINVAL;
    gw_run_graph* g = ctx->run_graph;
    if (!g || !g->exec) return gw_fail(ctx, GW_ESTATE, "gw_run_replay: nothing captured");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    // plan the run again: checks, allocations, region offsets of the current
    // walks; the graph is valid while nothing it points at moved
    RunPlan rp;
    int rc = plan_run(ctx, &g->plan.prog, rp);
    if (rc) return rc;
    if (fingerprint(ctx, rp) != g->fp)
        return gw_fail(ctx, GW_ESTATE, "gw_run_replay: buffers, graph or walk count changed since the capture");
    if ((rc = gen_table_ready(ctx, rp))) return rc;
    GW_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    GW_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    GW_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
    gw_stats st = g->plan.st0;
    st.total_ms = ms;
    st.step_ms = ms;  // the whole graph: reset, step kernel, aggregation, counter copy
    st.launches = ctx->nwalks ? 1 : 0;
    st.rounds = 1;
    if ((rc = finish_run(ctx, g->plan, g->h_cnt, st, 0.0))) return rc;
    ctx->finals_valid = !g->plan.L.no_final;
    if (stats) *stats = st;
    return GW_OK;
}

// Launch-shape autotuner: the same program at each (block_threads,
// waves_per_simd) candidate, best of `reps` runs after one warm-up; results
// do not depend on the launch shape (walks are keyed by id), only time does.
int gw_stream_policy(gw_ctx* ctx, uint32_t policy, double prob, uint64_t seed) {
    if (!ctx || policy > GW_BLOCKS_MAX_WEIGHT || !(prob >= 0.0 && prob <= 1.0))
        return gw_fail(ctx, GW_EINVAL, "gw_stream_policy: unknown policy or prob outside [0, 1]");
    ctx->block_policy = policy;
    ctx->block_prob = prob;
    ctx->block_seed = seed;
    return GW_OK;
}

int gw_autotune(gw_ctx* ctx, const gw_program* prog, uint32_t reps, gw_program* best, double* best_ms) {
    if (!ctx || !prog || !best || reps == 0) return gw_fail(ctx, GW_EINVAL, "gw_autotune: bad arguments");
    static const uint32_t kThreads[] = {64, 128, 256};
    static const uint32_t kWaves[] = {2, 4, 6, 8, 12, 16};
    const uint32_t sched = prog->schedule != GW_SCHED_AUTO ? prog->schedule
                           : ctx->host_graph            ? GW_SCHED_BLOCKED
                                                        : GW_SCHED_PERSISTENT;
    auto time_of = [&](const gw_program& p, double& ms) -> int {
        gw_stats st;
        int rc = gw_run(ctx, &p, &st);
        if (rc) return rc;
        ms = 1e300;
        for (uint32_t r = 0; r < reps; ++r) {
            if ((rc = gw_run(ctx, &p, &st))) return rc;
            ms = std::min(ms, st.total_ms);
        }
        return GW_OK;
    };
    // The blocked schedule on a device-resident graph: the block size first,
    // over uniform tables of 2^12 .. 2^24 vertices per block (at most 2048
    // blocks), at the program's launch shape; the fastest table stays
    // installed (trajectories do not depend on it).
    if (sched == GW_SCHED_BLOCKED && !ctx->host_graph && ctx->n > 0) {
        std::vector<uint32_t> best_table = ctx->block_first;
        double best_t = 1e300;
        for (uint32_t k = 12; k <= 24; k += 2) {
            const uint64_t bv = 1ull << k;
            const uint64_t nb = (ctx->n + bv - 1) / bv;
            if (nb > 2048) continue;
            std::vector<uint32_t> table(nb + 1);
            for (uint64_t b = 0; b < nb; ++b) table[b] = static_cast<uint32_t>(b * bv);
            table[nb] = ctx->n;
            int rc = gw_graph_blocks(ctx, table.data(), static_cast<uint32_t>(nb));
            double t = 0;
            if (!rc) rc = time_of(*prog, t);
            if (rc) return rc;
            if (t < best_t) {
                best_t = t;
                best_table = table;
            }
            if (nb == 1) break;  // larger blocks give the same one-block table
        }
        int rc = gw_graph_blocks(ctx, best_table.data(), static_cast<uint32_t>(best_table.size() - 1));
        if (rc) return rc;
    }
    double bms = 1e300;
    gw_program bp = *prog;
    for (uint32_t tb : kThreads) {
        for (uint32_t wps : kWaves) {
            if (sched == GW_SCHED_PLAIN && wps != kWaves[0]) continue;  // the plain grid has one walk per thread
            gw_program p = *prog;
            p.block_threads = tb;
            p.waves_per_simd = sched == GW_SCHED_PLAIN ? 0 : wps;
            double m = 0;
            int rc = time_of(p, m);
            if (rc) return rc;
            if (m < bms) {
                bms = m;
                bp = p;
            }
        }
    }
    // walks a wavefront claims per atomic, at the chosen shape (the plain
    // grid claims none)
    if (sched != GW_SCHED_PLAIN) {
        static const uint32_t kChunks[] = {128, 512, 2048, 8192};
        const gw_program shape = bp;
        for (uint32_t c : kChunks) {
            gw_program p = shape;
            p.walk_chunk = c;
            double m = 0;
            int rc = time_of(p, m);
            if (rc) return rc;
            if (m < bms) {
                bms = m;
                bp = p;
            }
        }
    }
    *best = bp;
    if (best_ms) *best_ms = bms;
    return GW_OK;
}

}  // extern "C"

// ---- gw_run_noalloc: the persistent schedule, allocating nothing ---------
//
// The run's device counters (steps, visits, the walk-claim cursor, skipped
// walks) live in a file-scope static array instead of the context's
// allocation.  Host code never passes the array's name to a kernel or a copy
// (in host code it is a host-side stand-in); it uses the device address
// hipGetSymbolAddress returns, looked up once per device by gw_create with
// that device current.  Each device holds its own copy; every call on one
// device shares it, so calls that overlap on one device race.
namespace gw {
namespace {

constexpr uint32_t kScratchWords = kCounters;  // 32 x u64 = 256 bytes per device
constexpr int kScratchDevices = 64;

std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kScratchDevices] = {};

// totals[0..2] = steps, visits, skipped walks of the run.
__global__ void noalloc_totals_kernel(const unsigned long long* __restrict__ cnt, unsigned long long* __restrict__ totals) {
    if (threadIdx.x == 0) {
        totals[0] = cnt[kCntSteps];
        totals[1] = cnt[kCntVisits];
        totals[2] = cnt[kCntBadWalks];
    }
}

template <int KIND>
hipError_t launch_noalloc(const Prog& p, const Graph& g, bool ext, bool vc, uint32_t grid, uint32_t tb, hipStream_t st,
                          const gw_walk* walks, uint64_t nw, uint32_t* counts, unsigned long long* cnt) {
    const EventLog log{nullptr, 0, 0};
    const SourceTable table{};
    const KeyLog klog{};
    constexpr int VC = GW_REC_VERTEX_COUNT;
    if (ext && vc)
        step_persistent_kernel<KIND, VC, true><<<dim3(grid), dim3(tb), 0, st>>>(p, g, walks, nullptr, nw, counts, nullptr,
                                                                            nullptr, cnt, log, table, klog);
    else if (ext)
        step_persistent_kernel<KIND, 0, true><<<dim3(grid), dim3(tb), 0, st>>>(p, g, walks, nullptr, nw, nullptr, nullptr,
                                                                           nullptr, cnt, log, table, klog);
    else if (vc)
        step_persistent_kernel<KIND, VC, false><<<dim3(grid), dim3(tb), 0, st>>>(p, g, walks, nullptr, nw, counts, nullptr,
                                                                             nullptr, cnt, log, table, klog);
    else
        step_persistent_kernel<KIND, 0, false><<<dim3(grid), dim3(tb), 0, st>>>(p, g, walks, nullptr, nw, nullptr, nullptr,
                                                                            nullptr, cnt, log, table, klog);
    return hipGetLastError();
}

}  // namespace
}  // namespace gw

static __device__ __align__(256) unsigned long long scratch[gw::kScratchWords];

int gw_noalloc_init(gw_ctx* ctx) {
    using namespace gw;
    if (ctx->device < 0 || ctx->device >= kScratchDevices)
        return gw_fail(ctx, GW_EINVAL, "gw_create: device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return gw_hip_check(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[ctx->device] = static_cast<unsigned long long*>(ptr);
    }
    ctx->d_scratch = g_scratch_ptr[ctx->device];
    return GW_OK;
}

extern "C" int gw_run_noalloc(gw_ctx* ctx, const gw_program* prog, const gw_walk* walks, uint64_t nwalks,
                              uint32_t* counts, unsigned long long* totals, void* stream) {
    using namespace gw;
    if (!ctx || !prog || !totals || (nwalks && !walks)) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: bad arguments");
    if (!ctx->d_scratch) return gw_fail(ctx, GW_ESTATE, "gw_run_noalloc: no scratch address for this device");
    if (!ctx->d_off) return gw_fail(ctx, GW_ESTATE, "gw_run_noalloc: needs a device-resident graph");
    if (nwalks > 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: at most 2^32 - 1 walks per call");
    if (prog->kind > GW_RESTART) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: unknown program kind");
    if (prog->L == 0 && prog->kind != GW_PROB) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: L must be >= 1");
    if (prog->stop_thr > (1ull << 32)) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: stop_thr above 2^32");
    if (prog->kind == GW_PROB && prog->stop_thr < (1ull << 20))
        return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: PROB needs stop_thr >= 2^20");
    if (prog->kind == GW_JUMP && (prog->nvertices == 0 || prog->nvertices > ctx->n))
        return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: JUMP needs nvertices in [1, n]");
    const uint32_t rec = prog->recorder & ~GW_REC_NO_FINAL;
    if (rec & ~GW_REC_VERTEX_COUNT)
        return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: records vertex counts only (GW_REC_VERTEX_COUNT or none)");
    const bool vc = rec == GW_REC_VERTEX_COUNT;
    if (vc != (counts != nullptr))
        return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: counts is required with GW_REC_VERTEX_COUNT and only then");
    if (prog->schedule != GW_SCHED_AUTO && prog->schedule != GW_SCHED_PERSISTENT)
        return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: runs the persistent schedule");
    // the successor-header layout only when an earlier run built it (building allocates)
    if (prog->layout == GW_LAYOUT_EXT && !ctx->d_ext)
        return gw_fail(ctx, GW_ESTATE, "gw_run_noalloc: successor-header edges not built (run gw_run with GW_LAYOUT_EXT first)");
    if (prog->layout > GW_LAYOUT_EXT) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: unknown layout");
    const bool ext = prog->layout != GW_LAYOUT_CSR && ctx->d_ext != nullptr;
    const uint32_t chunk = prog->walk_chunk ? prog->walk_chunk : kChunk;
    if (chunk % 64 != 0 || chunk > 8192)
        return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: walk_chunk must be a multiple of 64, at most 8192");
    const uint32_t tb = prog->block_threads ? prog->block_threads : 256u;
    if (tb % 64 != 0 || tb > 256) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: block_threads must be a multiple of 64, at most 256");
    const uint32_t wps = prog->waves_per_simd ? prog->waves_per_simd : 8u;
    if (wps > 32) return gw_fail(ctx, GW_EINVAL, "gw_run_noalloc: waves_per_simd must be at most 32");
    const uint64_t waves = static_cast<uint64_t>(std::max(1u, ctx->cu_count)) * 4u * wps;
    const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, waves * 64 / tb));
    GW_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    unsigned long long* cnt = ctx->d_scratch;
    // the counters start at zero; fill kernels (not memset nodes) keep the
    // sequence capturable into a HIP graph
    fill_u64_kernel<<<dim3(1), dim3(64), 0, st>>>(cnt, kScratchWords);
    GW_HIP(ctx, hipGetLastError());
    if (vc) {
        fill_u32_kernel<<<dim3(gw_grid(ctx->n, 256)), dim3(256), 0, st>>>(counts, ctx->n);
        GW_HIP(ctx, hipGetLastError());
    }
    if (nwalks) {
        const Prog p{prog->L, prog->nvertices, prog->stop_thr, static_cast<uint32_t>(prog->seed),
                     static_cast<uint32_t>(prog->seed >> 32), prog->stream, 0u, chunk};
        const Graph g{ctx->d_off, ctx->d_col, reinterpret_cast<const Edge*>(ctx->d_ext), ctx->n};
        hipError_t e;
        switch (prog->kind) {
        case GW_JUMP: e = launch_noalloc<GW_JUMP>(p, g, ext, vc, grid, tb, st, walks, nwalks, counts, cnt); break;
        case GW_STOP: e = launch_noalloc<GW_STOP>(p, g, ext, vc, grid, tb, st, walks, nwalks, counts, cnt); break;
        case GW_PROB: e = launch_noalloc<GW_PROB>(p, g, ext, vc, grid, tb, st, walks, nwalks, counts, cnt); break;
        case GW_RESTART_JOINT: e = launch_noalloc<GW_RESTART_JOINT>(p, g, ext, vc, grid, tb, st, walks, nwalks, counts, cnt); break;
        default: e = launch_noalloc<GW_RESTART>(p, g, ext, vc, grid, tb, st, walks, nwalks, counts, cnt); break;
        }
        if (e != hipSuccess) return gw_hip_check(ctx, e, "gw_run_noalloc: step kernel launch");
    }
    noalloc_totals_kernel<<<dim3(1), dim3(64), 0, st>>>(cnt, totals);
    GW_HIP(ctx, hipGetLastError());
    return GW_OK;
}
