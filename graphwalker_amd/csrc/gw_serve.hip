// gw_serve.hip -- continuous-batching query serving on the walk engine.
//
// Queries (source vertex, walk count) arrive at any time.  Admitted queries
// share one device pool of walk records; every serve step advances each live
// walk by at most `hops_per_step` hops in one persistent launch, so queries
// admitted later join walks already in flight, and a query completes when the
// last of its walks ends.  Per-query visit counts accumulate in a private
// open-addressing hash region of a device arena; admission is by free walk
// slots and arena space (the HBM budget), first come first served.  The
// random stream of a walk is keyed by (seed, query id; walk index, hop), so a
// query's result does not depend on what else is in flight.
//
// Over a host-resident graph (gw_graph_set_host, gw_graph_rmat_host) the
// offsets are copied to HBM and the columns stay in pinned host memory,
// read through a page table: pages of 2^k edges that the policy keeps in a
// device page cache (the graph's cache_bytes, offsets included), the others
// read from host memory directly.  Between serve steps the policy follows
// the kernel's sampled per-page demand and streams the most-read pages in
// with hipMemcpyAsync on the copy stream while steps run (the reference's
// block cache keeps the blocks with the most walks, graphwalker.hpp:168-251);
// no walk waits for a page.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <deque>
#include <map>
#include <numeric>
#include <set>
#include <vector>

#include "gw_device.h"
#include "gw_internal.h"

namespace gw {
namespace {

struct SlotInfo {
    uint64_t base;     // region offset in the arena (slots)
    uint32_t mask;     // region slots - 1
    uint32_t source;   // source vertex: start and restart target
    uint32_t qid_lo;   // Philox counter word 3
    uint32_t qid_hi;   // mixed into Philox key word 1
    uint64_t srow;     // the source's row and degree, written at admission:
    uint32_t sdeg;     // a restart needs no load
    uint32_t pad;
};

struct Admit;
struct Done;

struct ServeState {
    gw_program prog{};
    uint32_t hops = 0;
    uint64_t pool_cap = 0, arena_slots = 0, live = 0;
    uint32_t max_queries = 0, cbits = 0;
    gw_walk* pool[2] = {nullptr, nullptr};
    int cur = 0;
    unsigned long long* arena = nullptr;
    uint32_t* d_occ = nullptr;            // per arena slot: each region's claimed slots, in claim order
    uint32_t* d_occn = nullptr;           // per query slot: length of its occupied list
    SlotInfo* d_slots = nullptr;
    uint32_t* d_remaining = nullptr;
    uint32_t* d_srccount = nullptr;       // visits of each query to its own source
    bool ext = false;
    unsigned long long* d_cnt = nullptr;  // [0] out cursor [1] steps [2] visits [3] overflow [4] compact n
    unsigned long long* d_pairs = nullptr;  // compaction output
    uint64_t pairs_cap = 0;
    Done* d_done = nullptr;                 // max_queries completion records
    unsigned long long* d_region_n = nullptr;  // max_queries: entries of each completed query
    // per query slot: its key region in the arena (u32 index range) and cursor
    uint64_t* d_koff = nullptr;
    uint64_t* d_kend = nullptr;
    unsigned long long* d_kcur = nullptr;
    // arena range of an active query: [base, base + size) its hash region,
    // then `total - size` slots holding its key region (walks x L u32 keys)
    struct Active { uint64_t qid; uint32_t slot, walks; uint64_t base, size, total; uint32_t source; };
    // pinned host staging of the per-step copies (a copy from pageable memory
    // is staged by the runtime before the call returns)
    void* h_block = nullptr;
    Admit* h_adm = nullptr;
    Done* h_done = nullptr;
    unsigned long long* h_cnt = nullptr;   // 8 counters
    uint32_t* h_rem = nullptr;             // max_queries
    unsigned long long* h_kcur = nullptr;  // max_queries
    uint32_t* h_tn = nullptr;              // max_queries
    unsigned long long* h_top = nullptr;   // max_queries x topk (gw_serve_topk)
    // the device addresses of the pinned buffers the kernels read and write
    Admit* dh_adm = nullptr;
    unsigned long long* dh_cnt = nullptr;
    uint32_t* dh_rem = nullptr;
    unsigned long long* dh_kcur = nullptr;
    uint32_t* dh_tn = nullptr;
    unsigned long long* dh_top = nullptr;
    std::vector<Active> active;
    std::vector<uint32_t> free_slots;  // a min-heap: the lowest free slot is taken first
    std::vector<std::pair<uint64_t, uint64_t>> free_ranges;  // arena [start, size), sorted by start
    std::deque<gw_query> queue;
    std::map<uint64_t, std::vector<unsigned long long>> done;  // qid -> packed (vertex << cbits | count)
    std::deque<uint64_t> done_order;
    uint32_t topk = 0;                         // > 0: completed queries keep their top-k pairs only
    unsigned long long* d_cand = nullptr;      // per-chunk top-k candidates of the completing queries
    uint64_t cand_cap = 0;                     // keys
    std::set<uint64_t> qids;                   // queued, in flight, or completed and not yet read
    bool failed = false;                       // a step failed after its kernel ran: only close is allowed
    // host-resident graph: offsets in HBM, columns paged
    bool ooc = false;
    uint64_t* d_off = nullptr;                 // n + 1 offsets
    uint32_t page_log2 = 0;
    uint64_t npages = 0;
    const uint32_t* host_col = nullptr;        // device address of the pinned column array
    const uint32_t** h_pages = nullptr;        // pinned page table: rebased slot pointer or host_col
    const uint32_t** d_pages = nullptr;
    uint32_t* d_page_hits = nullptr;           // sampled column reads per page in the last step
    uint32_t* hits = nullptr;                  // pinned (coherent): written by each step's export kernel
    uint32_t* dh_hits = nullptr;               // its device address
    std::vector<double> score;                 // decayed demand
    std::vector<int32_t> page_slot;            // slot holding the page, -1 = read from host memory
    std::vector<int64_t> slot_page;            // page in (or loading into) the slot, -1 = free
    std::vector<char> loading;                 // per page
    uint32_t* d_cache = nullptr;               // slots x page edges
    uint64_t nslots = 0;
    uint32_t max_loads = 0;
    struct Load { uint32_t page, slot; hipEvent_t done; };
    std::vector<Load> inflight;
    std::vector<hipEvent_t> free_events;
    hipEvent_t table_ev = nullptr;
    bool table_dirty = true;
    uint64_t h2d_bytes = 0, pages_loaded = 0;
    uint32_t grid_max = 0;                     // serve-step workgroups resident on the device at once
};

// The context's serving state (null when no server is open on it).
ServeState* state_of(gw_ctx* ctx) { return static_cast<ServeState*>(ctx->serve_state); }

uint32_t bits_for(uint64_t x) {
    uint32_t b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

bool arena_alloc(ServeState& S, uint64_t size, uint64_t& base) {
    for (size_t i = 0; i < S.free_ranges.size(); ++i) {
        auto& r = S.free_ranges[i];
        if (r.second >= size) {
            base = r.first;
            r.first += size;
            r.second -= size;
            if (r.second == 0) S.free_ranges.erase(S.free_ranges.begin() + i);
            return true;
        }
    }
    return false;
}

void arena_free(ServeState& S, uint64_t base, uint64_t size) {
    auto it = std::lower_bound(S.free_ranges.begin(), S.free_ranges.end(), std::make_pair(base, uint64_t(0)));
    it = S.free_ranges.insert(it, {base, size});
    // coalesce with neighbours
    size_t i = it - S.free_ranges.begin();
    if (i + 1 < S.free_ranges.size() && S.free_ranges[i].first + S.free_ranges[i].second == S.free_ranges[i + 1].first) {
        S.free_ranges[i].second += S.free_ranges[i + 1].second;
        S.free_ranges.erase(S.free_ranges.begin() + i + 1);
    }
    if (i > 0 && S.free_ranges[i - 1].first + S.free_ranges[i - 1].second == S.free_ranges[i].first) {
        S.free_ranges[i - 1].second += S.free_ranges[i].second;
        S.free_ranges.erase(S.free_ranges.begin() + i);
    }
}

__device__ __forceinline__ uint32_t rank_of(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Per-visit recorder of a served walk: the step's visit is held, then staged
// by region_add_uniform after the step with the whole wavefront present
// (the batched path's direct regions, one field per query slot): a query's
// visits go to its key region in the arena through per-wavefront LDS buffers
// -- no atomic per visit -- and visits to the query's own source are only
// counted.  A completing query's keys are folded into its hash region by
// serve_fold_kernel.  Aggregating in the step instead (a per-wavefront LDS
// hash in front of the arena) put an LDS round trip and the combining of
// equal keys into every hop of every walk.
struct ServeVisits {
    uint32_t visits = 0, steps = 0;
    uint32_t pend_v = 0;
    bool pend = false;
    __device__ __forceinline__ void visit(uint32_t v, uint32_t) {
        ++visits;
        pend_v = v;
        pend = true;
    }
};

// One serve step: every walk of pool_in advances at most `hops` hops; walks
// still alive are appended to pool_out; ended walks decrement their query's
// remaining count.  Persistent wavefronts take chunks of the pool.
// Query slot q's visits (other than to its source) go to its key region
// keys[koff[q], kend[q]) through the cursor kcur[q]; visits to its source
// count in srccount[q].
template <int KIND, bool EXT, bool PAGED>
__global__ __launch_bounds__(256) void serve_step_kernel(Prog p, Graph g, PagedCols pc, const gw_walk* __restrict__ pool_in,
                                                        uint64_t n, gw_walk* __restrict__ pool_out,
                                                        const SlotInfo* __restrict__ slots,
                                                        uint32_t* __restrict__ remaining,
                                                        uint32_t* __restrict__ srccount,
                                                        uint32_t* __restrict__ keys,
                                                        const uint64_t* __restrict__ koff,
                                                        const uint64_t* __restrict__ kend,
                                                        unsigned long long* __restrict__ kcur, uint32_t nslots,
                                                        uint32_t hops,
                                                        unsigned long long* __restrict__ cnt,
                                                        unsigned long long* __restrict__ cursor, uint32_t chunk,
                                                        uint64_t nchunks, uint64_t stride) {
    const uint32_t lane = __lane_id();
    uint64_t chunk_next = 0, chunk_end = 0;
    bool exhausted = false, active = false;
    uint32_t slot = 0, budget = 0, tot_steps = 0, tot_visits = 0;
    Prog q = p;
    KeyLog kl{};
    kl.region_off = koff;
    kl.region_end = kend;
    kl.region_n = kcur;
    kl.region_out = keys;
    kl.region_overflow = cnt + 3;
    kl.nf = nslots;
    kl.field_count = srccount;
    __shared__ RegionBuf s_rbuf[4];
    RegionBuf* rbuf = &s_rbuf[threadIdx.x >> 6];
    RegionState rst;
    ServeVisits rec;
    State s{};
    for (;;) {
        uint64_t idle = __ballot(!active);
        while (idle && !exhausted) {
            if (chunk_next == chunk_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(cursor, 1ull);
                base = lane_value(base, 0u);
                if (base >= nchunks) { exhausted = true; break; }
                // chunks in a permuted order (stride coprime to nchunks): the
                // pool holds each query's walks back to back, and taken in
                // order the chunks in flight would come from a few large
                // queries, whose walks all read the same few lines
                base = (base * stride % nchunks) * chunk;  // both below nchunks < 2^32
                chunk_next = base;
                chunk_end = base + chunk < n ? base + chunk : n;
            }
            const uint32_t want = __popcll(idle);
            const uint64_t avail = chunk_end - chunk_next;
            const uint32_t take = avail < want ? static_cast<uint32_t>(avail) : want;
            const uint32_t r = rank_of(idle);
            if (!active && r < take) {
                const gw_walk w = pool_in[chunk_next + r];
                slot = w.source;
                const SlotInfo si = slots[slot];
                s.id = w.id;
                s.source = si.source;
                s.v = w.vertex;
                s.hop = w.hop;
                load_row<EXT>(g, s);
                s.srow = si.srow;
                s.sdeg = si.sdeg;
                q.stream = si.qid_lo;
                q.k1 = p.k1 ^ si.qid_hi;
                start_walk(q, s);
                rec.visits = 0;
                rec.steps = 0;
                budget = hops;
                active = true;
            }
            chunk_next += take;
            idle = __ballot(!active);
        }
        if (__ballot(active) == 0) break;
        bool leave = false, done = false;
        if (active) {
            const uint32_t hop_before = s.hop;
            if constexpr (PAGED)
                done = step_once<KIND, false>(q, g, s, rec, pc);
            else
                done = step_once<KIND, EXT>(q, g, s, rec);
            budget -= (s.hop != hop_before) ? 1u : 0u;
            leave = !done && budget == 0;
        }
        // the step's visit into its query's key region, every lane present
        // (the query's source is the field's counted-apart vertex)
        region_add_uniform<true>(kl, rbuf, rst, rec.pend, slot, rec.pend_v, s.source);
        rec.pend = false;
        // walks that ended: one decrement per (wavefront, query)
        uint64_t todo = __ballot(done);
        while (todo) {
            const uint32_t leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
            const uint32_t ls = lane_value(slot, leader);
            const uint64_t grp = __ballot(done && slot == ls);
            if (lane == leader) atomicSub(&remaining[ls], static_cast<uint32_t>(__popcll(grp)));
            todo &= ~grp;
        }
        if (active) {
            if (done || leave) {
                tot_steps += rec.steps;
                tot_visits += rec.visits;
                active = false;
            }
        }
        // survivors go to the next pool: one atomic per wavefront
        const uint64_t lm = __ballot(leave);
        if (lm) {
            const uint32_t leader = __ffsll(static_cast<unsigned long long>(lm)) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&cnt[0], static_cast<unsigned long long>(__popcll(lm)));
            base = lane_value(base, leader);
            if (leave) pool_out[base + rank_of(lm)] = gw_walk{s.id, slot, s.v, s.hop};
        }
    }
    region_evict(kl, rbuf->key[0], rst.f0, rst.c0, rst.d0);  // every lane of the wavefront is here
    region_evict(kl, rbuf->key[1], rst.f1, rst.c1, rst.d1);
    workgroup_add2(&cnt[1], &cnt[2], tot_steps, tot_visits);
}

// A query admitted in this step: its slot record, walk count and where its
// walks start in the pool.
struct Admit {
    SlotInfo info;
    uint64_t at;
    uint32_t slot, walks;
    uint64_t koff, kend;  // its key region, as u32 indexes into the arena
};

// One workgroup per admitted query: writes its slot record and remaining
// count, then its walk records (id = walk index, source field = slot).
// Reads the admission records from pinned host memory (no copy to enqueue).
__global__ void serve_admit_kernel(const Admit* __restrict__ adm, gw_walk* __restrict__ pool,
                                   SlotInfo* __restrict__ slots, uint32_t* __restrict__ remaining,
                                   uint32_t* __restrict__ srccount, uint32_t* __restrict__ occn,
                                   const uint64_t* __restrict__ off, uint64_t* __restrict__ koff,
                                   uint64_t* __restrict__ kend, unsigned long long* __restrict__ kcur) {
    const Admit a = adm[blockIdx.x];
    if (threadIdx.x == 0) {
        SlotInfo si = a.info;
        si.srow = off[si.source];
        si.sdeg = static_cast<uint32_t>(off[si.source + 1] - si.srow);
        slots[a.slot] = si;
        koff[a.slot] = a.koff;
        kend[a.slot] = a.kend;
        kcur[a.slot] = 0;
        remaining[a.slot] = a.walks;
        srccount[a.slot] = 0;
        occn[a.slot] = 0;
    }
    for (uint32_t i = threadIdx.x; i < a.walks; i += blockDim.x)
        pool[a.at + i] = gw_walk{i, a.slot, a.info.source, 0u};
}

// After the walks of a serve step: the counters, and the remaining walks
// and key cursors of the slots below `hi`, into pinned host memory (one
// launch in place of a copy each); the 8 counters are then zeroed for the
// next step (the fold adds its overflow count to counter 4 afterwards).
__global__ __launch_bounds__(256) void serve_export_kernel(unsigned long long* __restrict__ cnt,
                                                           const uint32_t* __restrict__ rem,
                                                           const unsigned long long* __restrict__ kcur, uint32_t hi,
                                                           unsigned long long* __restrict__ o_cnt,
                                                           uint32_t* __restrict__ o_rem,
                                                           unsigned long long* __restrict__ o_kcur,
                                                           uint32_t* __restrict__ hits, uint32_t* __restrict__ o_hits,
                                                           uint32_t npages) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    if (i < 4) o_cnt[i] = cnt[i];
    if (i < 8) cnt[i] = 0;
    for (uint32_t j = i; j < hi; j += stride) {
        o_rem[j] = rem[j];
        o_kcur[j] = kcur[j];
    }
    // host-resident graph: the step's sampled reads per page, then cleared
    for (uint32_t j = i; j < npages; j += stride) {
        o_hits[j] = hits[j];
        hits[j] = 0;
    }
}

// A completed query's count region and where its pairs go.
struct Done {
    uint64_t base, size;   // hash region
    uint64_t out;          // first output index: room for nkeys pairs from there
    uint32_t slot, source; // the query's slot (its source-vertex counter) and source
    uint64_t kbase, nkeys; // its key region (u32 index into the arena) and the keys in it
    uint64_t cand;         // top-k: its first chunk candidate (the chunks' candidates back to back)
};

// A completed query's keys, folded into its hash region: one workgroup per
// kFoldKeys keys counts them in an LDS table, then adds each distinct vertex
// once to the hash region (grid.x = chunk, grid.y = query).  The keys are
// zeroed as they are read, so the arena range is clear again for the next
// query once the gather has cleared the hash slots.
constexpr uint32_t kFoldKeys = 4096;
constexpr uint32_t kFoldSlots = 8192;  // load at most one half
constexpr uint32_t kFoldEmpty = 0xFFFFFFFFu;  // never a vertex (vertex ids are below 2^32 - 1)

__global__ __launch_bounds__(512) void serve_fold_kernel(unsigned long long* __restrict__ arena,
                                                        uint32_t* __restrict__ occ, uint32_t* __restrict__ occn,
                                                        const Done* __restrict__ done,
                                                        const unsigned long long* __restrict__ kcur, uint32_t cbits,
                                                        unsigned long long* __restrict__ overflow) {
    __shared__ uint32_t tkey[kFoldSlots];
    __shared__ uint32_t tcnt[kFoldSlots];
    __shared__ uint32_t s_n, s_claims, s_at;
    const Done d = done[blockIdx.y];
    // the keys written: the query's key cursor, within the record's bound
    const uint64_t nk = std::min<uint64_t>(kcur[d.slot], d.nkeys);
    const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * kFoldKeys;
    if (lo >= nk) return;  // workgroup-uniform
    const uint64_t hi = std::min<uint64_t>(lo + kFoldKeys, nk);
    for (uint32_t i = threadIdx.x; i < kFoldSlots; i += blockDim.x) {
        tkey[i] = kFoldEmpty;
        tcnt[i] = 0;
    }
    if (threadIdx.x == 0) s_n = s_claims = s_at = 0;
    __syncthreads();
    uint32_t* keys = reinterpret_cast<uint32_t*>(arena) + d.kbase;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint32_t v = keys[i];
        keys[i] = 0;
        uint32_t h = static_cast<uint32_t>((static_cast<uint64_t>(v) * 0x9E3779B97F4A7C15ull) >> 51);
        for (;;) {  // at most kFoldKeys keys in kFoldSlots slots: an empty slot exists
            const uint32_t old = atomicCAS(&tkey[h], kFoldEmpty, v);
            if (old == kFoldEmpty || old == v) {
                atomicAdd(&tcnt[h], 1u);
                break;
            }
            h = (h + 1) & (kFoldSlots - 1);
        }
    }
    __syncthreads();
    // the distinct vertices, compacted to the front of the table (one LDS
    // atomic per wavefront): every thread reads its slots into registers
    // before any thread writes (the barrier between)
    uint32_t mv[kFoldSlots / 512], mc[kFoldSlots / 512];
#pragma unroll
    for (uint32_t j = 0; j < kFoldSlots / 512; ++j) {
        mv[j] = tkey[j * 512 + threadIdx.x];
        mc[j] = tcnt[j * 512 + threadIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kFoldSlots / 512; ++j) {
        const bool have = mv[j] != kFoldEmpty;
        const uint64_t bm = __ballot(have);
        uint32_t at = 0;
        if (bm) {
            const uint32_t leader = __ffsll(static_cast<unsigned long long>(bm)) - 1;
            if (__lane_id() == leader) at = atomicAdd(&s_n, static_cast<uint32_t>(__popcll(bm)));
            at = lane_value(at, leader);
        }
        if (have) {
            tkey[at + rank_of(bm)] = mv[j];
            tcnt[at + rank_of(bm)] = mc[j];
        }
    }
    __syncthreads();
    const uint32_t nd = s_n;
    // each distinct vertex into the hash region: a thread's compare-and-swaps
    // are issued back to back, then resolved; collisions probe on
    const uint32_t rm = static_cast<uint32_t>(d.size - 1);
    constexpr uint32_t kPer = kFoldKeys / 512;  // distinct vertices per thread at most
    uint32_t cv[kPer], cc[kPer], ch[kPer];
    unsigned long long cur[kPer];
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t i = j * 512 + threadIdx.x;
        cv[j] = i < nd ? tkey[i] : kFoldEmpty;
        cc[j] = i < nd ? tcnt[i] : 0u;
        ch[j] = static_cast<uint32_t>((static_cast<uint64_t>(cv[j]) * 0x9E3779B97F4A7C15ull) >> 32) & rm;
        cur[j] = cv[j] != kFoldEmpty
                     ? atomicCAS(arena + d.base + ch[j], 0ull, (static_cast<unsigned long long>(cv[j]) << cbits) | cc[j])
                     : 1ull;
    }
    uint32_t claimed_mask = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        if (cv[j] == kFoldEmpty) continue;
        const unsigned long long entry = (static_cast<unsigned long long>(cv[j]) << cbits) | cc[j];
        unsigned long long c = cur[j];
        uint32_t h = ch[j];
        bool placed = false;
        for (uint32_t probe = 0; probe <= rm; ++probe) {
            if (c == 0ull) {
                claimed_mask |= 1u << j;
                placed = true;
                break;
            }
            if ((c >> cbits) == cv[j]) {
                atomicAdd(arena + d.base + h, static_cast<unsigned long long>(cc[j]));
                placed = true;
                break;
            }
            h = (h + 1) & rm;
            c = atomicCAS(arena + d.base + h, 0ull, entry);
        }
        ch[j] = h;
        if (!placed) atomicAdd(overflow, static_cast<unsigned long long>(cc[j]));
    }
    // claimed slots join the query's occupied list: one global atomic per workgroup
    uint32_t mine = __popc(claimed_mask), rank0 = 0;
    if (mine) rank0 = atomicAdd(&s_claims, mine);
    __syncthreads();
    if (threadIdx.x == 0 && s_claims) s_at = atomicAdd(&occn[d.slot], s_claims);
    __syncthreads();
    const uint32_t base = s_at;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j)
        if (claimed_mask & (1u << j)) occ[d.base + base + rank0++] = ch[j];
}

// A completed query's entries: its occupied list names the hash slots the
// fold claimed; each entry is copied to the query's output range and its
// slot cleared for reuse (grid.y = query).  Writes each query's pair count
// to n[q] for the top-k kernels.
__global__ void serve_gather_kernel(unsigned long long* __restrict__ arena, const uint32_t* __restrict__ occ,
                                    const uint32_t* __restrict__ occn, const Done* __restrict__ done,
                                    unsigned long long* __restrict__ n, unsigned long long* __restrict__ out) {
    const Done d = done[blockIdx.y];
    const uint64_t m = occn[d.slot];
    if (blockIdx.x == 0 && threadIdx.x == 0) n[blockIdx.y] = m;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t h = occ[d.base + i];
        out[d.out + i] = arena[d.base + h];
        arena[d.base + h] = 0ull;
    }
}

// Top-k of completed queries (gw_serve_topk), in two levels so a query with
// many pairs spreads over many CUs.  Items of query q: its compacted pairs
// pairs[d.out .. d.out + n) (n = per-query pair count) and, as item n, its
// source vertex (counted apart, in srccount).  Level 1: one workgroup per
// kServeTopkChunk items writes that chunk's top-k keys to k candidate slots
// (empty slots hold key 0).  Level 2: one workgroup per query selects the
// top-k of its chunks' candidates -- the top-k of the union of every chunk's
// top-k is the query's top-k -- and writes them packed (vertex << cbits) |
// count, count descending, ties to the smaller vertex.
constexpr uint64_t kServeTopkChunk = 8192;

__device__ __forceinline__ uint64_t topk_chunks(uint64_t n) { return (n + 1 + kServeTopkChunk - 1) / kServeTopkChunk; }

__global__ __launch_bounds__(1024) void serve_topk_chunk_kernel(const unsigned long long* __restrict__ pairs,
                                                              const Done* __restrict__ done,
                                                              const unsigned long long* __restrict__ per_region,
                                                              const uint32_t* __restrict__ srccount,
                                                              uint32_t cbits, uint32_t k,
                                                              unsigned long long* __restrict__ cand) {
    __shared__ TopkScratch t;
    const uint32_t q = blockIdx.y;
    const uint64_t n = per_region[q];
    const uint64_t c = blockIdx.x;
    if (c >= topk_chunks(n)) return;  // workgroup-uniform
    const Done d = done[q];
    const unsigned long long cmask = (1ull << cbits) - 1;
    const uint32_t sc = srccount[d.slot];
    const uint64_t lo = c * kServeTopkChunk;
    const uint64_t cnt = std::min<uint64_t>(kServeTopkChunk, n + 1 - lo);
    // a chunk's candidates need no order: the merge selects and sorts
    const uint32_t m = block_topk<false, kServeTopkChunk / 1024>(
        cnt, k,
        [&](uint64_t i) {
            const uint64_t j = lo + i;
            if (j == n) return topk_key(d.source, sc);
            const unsigned long long x = pairs[d.out + j];
            return topk_key(static_cast<uint32_t>(x >> cbits), static_cast<uint32_t>(x & cmask));
        },
        t);
    unsigned long long* o = cand + d.cand + c * k;
    for (uint32_t r = threadIdx.x; r < k; r += blockDim.x) o[r] = r < m ? t.sel[r] : 0ull;
}

// The results go straight into pinned host memory (out, out_n, *fold_copy):
// the host reads them after its synchronisation, with no copy to enqueue.
__global__ __launch_bounds__(1024) void serve_topk_merge_kernel(const unsigned long long* __restrict__ per_region,
                                                              const Done* __restrict__ done,
                                                              const unsigned long long* __restrict__ cand,
                                                              uint32_t cbits, uint32_t k,
                                                              unsigned long long* __restrict__ out,
                                                              uint32_t* __restrict__ out_n,
                                                              const unsigned long long* __restrict__ fold_over,
                                                              unsigned long long* __restrict__ fold_copy) {
    __shared__ TopkScratch t;
    const uint32_t q = blockIdx.x;
    // the fold ran before this kernel: its overflow count travels with the results
    if (q == 0 && threadIdx.x == 0) *fold_copy = *fold_over;
    const unsigned long long* c = cand + done[q].cand;
    const uint32_t m = block_topk<true, 16>(topk_chunks(per_region[q]) * k, k, [&](uint64_t i) { return c[i]; }, t);
    for (uint32_t r = threadIdx.x; r < m; r += blockDim.x) {
        const unsigned long long key = t.sel[r];
        out[static_cast<uint64_t>(q) * k + r] =
            (static_cast<unsigned long long>(~static_cast<uint32_t>(key)) << cbits) | (key >> 32);
    }
    if (threadIdx.x == 0) out_n[q] = m;
}

template <int KIND>
hipError_t launch_serve(gw_ctx* ctx, ServeState& S, const Prog& p, uint64_t n, uint32_t grid) {
    const Graph g = S.ooc ? Graph{S.d_off, nullptr, nullptr, ctx->n}
                          : Graph{ctx->d_off, ctx->d_col, reinterpret_cast<const Edge*>(ctx->d_ext), ctx->n};
    const PagedCols pc{S.d_pages, S.d_page_hits, S.page_log2};
    // Walks claimed per wavefront at a time: while the pool fits the grid at
    // one walk per lane, 64, so every walk starts at once and a step lasts one
    // walk's hops; larger pools take up to 256 (fewer cursor atomics).
    const uint64_t waves = static_cast<uint64_t>(grid) * 4u;
    const uint64_t per_wave = (n + waves - 1) / waves;
    const uint32_t chunk = static_cast<uint32_t>(std::min<uint64_t>(256, std::max<uint64_t>(64, (per_wave + 63) / 64 * 64)));
    const uint64_t nchunks = (n + chunk - 1) / chunk;
    uint64_t stride = std::max<uint64_t>(1, static_cast<uint64_t>(static_cast<double>(nchunks) * 0.6180339887) | 1);
    while (std::gcd(stride, nchunks) != 1) stride += 2;
    stride %= nchunks;
    uint32_t* keys = reinterpret_cast<uint32_t*>(S.arena);
    if (S.ooc)
        serve_step_kernel<KIND, false, true><<<dim3(grid), dim3(256), 0, ctx->stream>>>(
            p, g, pc, S.pool[S.cur], n, S.pool[S.cur ^ 1], S.d_slots, S.d_remaining, S.d_srccount, keys, S.d_koff,
            S.d_kend, S.d_kcur, S.max_queries, S.hops, S.d_cnt, S.d_cnt + 5, chunk, nchunks, stride);
    else if (S.ext)
        serve_step_kernel<KIND, true, false><<<dim3(grid), dim3(256), 0, ctx->stream>>>(
            p, g, pc, S.pool[S.cur], n, S.pool[S.cur ^ 1], S.d_slots, S.d_remaining, S.d_srccount, keys, S.d_koff,
            S.d_kend, S.d_kcur, S.max_queries, S.hops, S.d_cnt, S.d_cnt + 5, chunk, nchunks, stride);
    else
        serve_step_kernel<KIND, false, false><<<dim3(grid), dim3(256), 0, ctx->stream>>>(
            p, g, pc, S.pool[S.cur], n, S.pool[S.cur ^ 1], S.d_slots, S.d_remaining, S.d_srccount, keys, S.d_koff,
            S.d_kend, S.d_kcur, S.max_queries, S.hops, S.d_cnt, S.d_cnt + 5, chunk, nchunks, stride);
    return hipGetLastError();
}

// Count-region slots of a query: twice its most distinct vertices, rounded
// up to a power of two.
uint64_t region_slots(uint64_t walks, uint32_t L, uint32_t n) {
    const uint64_t distinct = std::min<uint64_t>(walks * L, n);
    uint64_t size = 64;
    while (size < 2 * distinct) size <<= 1;
    return size;
}

// Arena slots of a query: its hash region, then its key region (a visit per
// hop of every walk at most, two u32 keys per slot).
uint64_t query_slots(uint64_t walks, uint32_t L, uint32_t n) {
    return region_slots(walks, L, n) + (walks * L + 1) / 2;
}

// ---- page cache of a server over a host-resident graph ------------------

const uint32_t* page_base(const ServeState& S, uint64_t slot, uint64_t page) {
    // rebased so that base[e] is column e for e in the page
    const uintptr_t slot_addr = reinterpret_cast<uintptr_t>(S.d_cache) + (slot << S.page_log2) * 4;
    return reinterpret_cast<const uint32_t*>(slot_addr - (page << S.page_log2) * 4);
}

uint64_t page_edges(const gw_ctx* ctx, const ServeState& S, uint64_t page) {
    const uint64_t lo = page << S.page_log2;
    return std::min<uint64_t>(1ull << S.page_log2, ctx->m - lo);
}

int ooc_open(gw_ctx* ctx, ServeState& S) {
    S.ooc = true;
    int unified = 0;
    GW_HIP(ctx, hipDeviceGetAttribute(&unified, hipDeviceAttributeUnifiedAddressing, ctx->device));
    if (!unified) return gw_fail(ctx, GW_EHIP, "gw_serve_open: the device cannot read pinned host memory in place");
    const uint64_t off_bytes = (ctx->n + 1ull) * 8;
    if (ctx->cache_bytes < off_bytes)
        return gw_fail(ctx, GW_ENOMEM, "gw_serve_open: the graph's cache_bytes (" + std::to_string(ctx->cache_bytes) +
                                           ") does not hold its offsets (" + std::to_string(off_bytes) + " bytes)");
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&S.d_off, off_bytes, "serve offsets"))) return rc;
    GW_HIP(ctx, hipMemcpy(S.d_off, ctx->h_off, off_bytes, hipMemcpyHostToDevice));
    uint32_t lg = ctx->serve_page_log2;
    if (lg == 0) {  // pages of at least 2^10 edges, at most 4096 pages
        lg = 10;
        while (lg < 26 && (ctx->m >> lg) >= 4096) ++lg;
    }
    S.page_log2 = lg;
    S.npages = (ctx->m + (1ull << lg) - 1) >> lg;
    const uint64_t page_bytes = (1ull << lg) * 4;
    S.nslots = std::min<uint64_t>(S.npages, (ctx->cache_bytes - off_bytes) / page_bytes);
    S.max_loads = ctx->serve_max_loads ? ctx->serve_max_loads : 4u;
    S.host_col = ctx->h_col;  // pinned memory, read in place over the host link
    const uint64_t np = std::max<uint64_t>(1, S.npages);
    GW_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&S.h_pages), np * sizeof(void*), hipHostMallocDefault));
    if ((rc = gw_alloc(ctx, (void**)&S.d_pages, np * sizeof(void*), "serve page table")) ||
        (rc = gw_alloc(ctx, (void**)&S.d_page_hits, np * 4, "serve page demand")))
        return rc;
    if (S.nslots && (rc = gw_alloc(ctx, (void**)&S.d_cache, S.nslots * page_bytes, "serve page cache"))) return rc;
    if (hipHostMalloc(reinterpret_cast<void**>(&S.hits), np * 4, hipHostMallocCoherent) != hipSuccess) {
        S.hits = nullptr;
        return gw_fail(ctx, GW_ENOMEM, "gw_serve_open: pinned page demand buffer");
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&S.dh_hits), S.hits, 0) != hipSuccess || !S.dh_hits)
        return gw_fail(ctx, GW_ENOMEM, "gw_serve_open: device address of the page demand buffer");
    std::fill(S.hits, S.hits + np, 0u);
    S.score.assign(S.npages, 0.0);
    S.page_slot.assign(S.npages, -1);
    S.loading.assign(S.npages, 0);
    S.slot_page.assign(S.nslots, -1);
    // the cache starts with the first pages (the rows of the lowest vertex
    // ids); demand decides from the first step on
    for (uint64_t pg = 0; pg < S.npages; ++pg) S.h_pages[pg] = S.host_col;
    for (uint64_t sl = 0; sl < S.nslots; ++sl) {
        GW_HIP(ctx, hipMemcpy(S.d_cache + (sl << lg), ctx->h_col + (sl << lg), page_edges(ctx, S, sl) * 4,
                              hipMemcpyHostToDevice));
        S.page_slot[sl] = static_cast<int32_t>(sl);
        S.slot_page[sl] = static_cast<int64_t>(sl);
        S.h_pages[sl] = page_base(S, sl, sl);
    }
    GW_HIP(ctx, hipMemcpy(S.d_pages, S.h_pages, np * sizeof(void*), hipMemcpyHostToDevice));
    GW_HIP(ctx, hipMemset(S.d_page_hits, 0, np * 4));
    S.table_dirty = false;
    if (!ctx->copy_stream) GW_HIP(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    GW_HIP(ctx, hipEventCreateWithFlags(&S.table_ev, hipEventDisableTiming));
    return GW_OK;
}

// Before a serve step's kernel (the engine stream is idle): loads that
// finished enter the table; demand decays and the most-read pages that are
// read from host memory replace the least-read resident ones (or fill free
// slots), at most max_loads copies in flight; the table goes to the device
// and the copies start once it is there, so no kernel reads a slot while it
// is overwritten.  *started = bytes of the loads started.
int ooc_before_step(gw_ctx* ctx, ServeState& S, uint64_t* started, uint32_t* nstarted) {
    *started = 0;
    *nstarted = 0;
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < S.inflight.size();) {
        const hipError_t q = hipEventQuery(S.inflight[i].done);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        if (q != hipSuccess) return gw_hip_check(ctx, q, "serve page load");
        const ServeState::Load ld = S.inflight[i];
        S.page_slot[ld.page] = static_cast<int32_t>(ld.slot);
        S.loading[ld.page] = 0;
        S.h_pages[ld.page] = page_base(S, ld.slot, ld.page);
        S.table_dirty = true;
        S.free_events.push_back(ld.done);
        S.inflight.erase(S.inflight.begin() + i);
    }
    for (uint64_t pg = 0; pg < S.npages; ++pg) S.score[pg] = S.score[pg] * 0.75 + S.hits[pg];
    std::vector<ServeState::Load> start;
    if (S.nslots && S.inflight.size() < S.max_loads) {
        std::vector<uint32_t> cand, res;
        for (uint64_t pg = 0; pg < S.npages; ++pg) {
            if (S.loading[pg]) continue;
            if (S.page_slot[pg] < 0) {
                if (S.score[pg] > 0) cand.push_back(static_cast<uint32_t>(pg));
            } else {
                res.push_back(static_cast<uint32_t>(pg));
            }
        }
        const size_t want = std::min<size_t>(cand.size(), S.max_loads - S.inflight.size());
        std::partial_sort(cand.begin(), cand.begin() + want, cand.end(),
                          [&](uint32_t a, uint32_t b) { return S.score[a] > S.score[b]; });
        std::sort(res.begin(), res.end(), [&](uint32_t a, uint32_t b) { return S.score[a] < S.score[b]; });
        std::vector<uint64_t> free_slots;
        for (uint64_t sl = 0; sl < S.nslots; ++sl)
            if (S.slot_page[sl] < 0) free_slots.push_back(sl);
        size_t vi = 0;
        for (size_t ci = 0; ci < want; ++ci) {
            const uint32_t c = cand[ci];
            uint64_t slot;
            if (!free_slots.empty()) {
                slot = free_slots.back();
                free_slots.pop_back();
            } else {
                if (vi >= res.size()) break;
                const uint32_t v = res[vi];
                if (S.score[c] <= 1.5 * S.score[v] + 1.0) break;  // hysteresis: no swap for a small gain
                ++vi;
                slot = static_cast<uint64_t>(S.page_slot[v]);
                S.page_slot[v] = -1;
                S.h_pages[v] = S.host_col;
                S.table_dirty = true;
            }
            S.slot_page[slot] = c;
            S.loading[c] = 1;
            start.push_back({c, static_cast<uint32_t>(slot), nullptr});
        }
    }
    if (S.table_dirty) {
        GW_HIP(ctx, hipMemcpyAsync(S.d_pages, S.h_pages, S.npages * sizeof(void*), hipMemcpyHostToDevice, ctx->stream));
        S.table_dirty = false;
    }
    if (!start.empty()) {
        GW_HIP(ctx, hipEventRecord(S.table_ev, ctx->stream));
        GW_HIP(ctx, hipStreamWaitEvent(ctx->copy_stream, S.table_ev, 0));
        for (auto& ld : start) {
            if (S.free_events.empty()) {
                hipEvent_t ev = nullptr;
                GW_HIP(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                S.free_events.push_back(ev);
            }
            ld.done = S.free_events.back();
            S.free_events.pop_back();
            const uint64_t bytes = page_edges(ctx, S, ld.page) * 4;
            GW_HIP(ctx, hipMemcpyAsync(S.d_cache + (static_cast<uint64_t>(ld.slot) << S.page_log2),
                                       ctx->h_col + (static_cast<uint64_t>(ld.page) << S.page_log2), bytes,
                                       hipMemcpyHostToDevice, ctx->copy_stream));
            GW_HIP(ctx, hipEventRecord(ld.done, ctx->copy_stream));
            S.inflight.push_back(ld);
            *started += bytes;
            ++*nstarted;
        }
        S.h2d_bytes += *started;
        S.pages_loaded += *nstarted;
    }
    // the demand counts were read out and cleared by the last step's export kernel
    return GW_OK;
}

void ooc_close(gw_ctx* ctx, ServeState& S) {
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    for (auto& ld : S.inflight) (void)hipEventDestroy(ld.done);
    for (auto ev : S.free_events) (void)hipEventDestroy(ev);
    S.inflight.clear();
    S.free_events.clear();
    if (S.table_ev) (void)hipEventDestroy(S.table_ev);
    if (S.h_pages) (void)hipHostFree(S.h_pages);
    if (S.hits) (void)hipHostFree(S.hits);
    S.hits = S.dh_hits = nullptr;
    const uint64_t np = std::max<uint64_t>(1, S.npages);
    gw_free(ctx, S.d_pages, np * sizeof(void*));
    gw_free(ctx, S.d_page_hits, np * 4);
    gw_free(ctx, S.d_cache, S.nslots * (1ull << S.page_log2) * 4);
    gw_free(ctx, S.d_off, (ctx->n + 1ull) * 8);
}

}  // namespace
}  // namespace gw

using namespace gw;

static void gw_serve_close_state(gw_ctx* ctx, gw::ServeState* S);

extern "C" {

int gw_serve_open(gw_ctx* ctx, const gw_program* prog, uint64_t pool_walks, uint64_t arena_bytes,
                  uint32_t max_queries, uint32_t hops_per_step) {
    if (!ctx || !prog || pool_walks == 0 || max_queries == 0 || hops_per_step == 0 || arena_bytes < 4096)
        return gw_fail(ctx, GW_EINVAL, "gw_serve_open: bad arguments");
    if (max_queries > 65535)  // completion launches one grid row per query
        return gw_fail(ctx, GW_EINVAL, "gw_serve_open: max_queries must be at most 65535");
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_serve_open: no graph");
    if (ctx->host_graph && prog->layout == GW_LAYOUT_EXT)
        return gw_fail(ctx, GW_EINVAL, "gw_serve_open: a host-resident graph is read in the CSR layout");
    if (prog->kind != GW_RESTART && prog->kind != GW_STOP && prog->kind != GW_JUMP)
        return gw_fail(ctx, GW_EINVAL, "gw_serve_open: serving runs RESTART, STOP or JUMP programs");
    if (prog->L == 0 || prog->stop_thr > (1ull << 32)) return gw_fail(ctx, GW_EINVAL, "gw_serve_open: bad program");
    if (state_of(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_serve_open: already open");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    ServeState* S = new ServeState();
    S->prog = *prog;
    S->prog.nvertices = prog->kind == GW_JUMP ? (prog->nvertices ? prog->nvertices : ctx->n) : prog->nvertices;
    S->hops = hops_per_step;
    S->pool_cap = pool_walks;
    S->arena_slots = arena_bytes / 8;
    S->max_queries = max_queries;
    S->cbits = 64 - std::max(1u, bits_for(ctx->n ? ctx->n - 1ull : 0ull));
    int rc = 0;
    if ((rc = gw_alloc(ctx, (void**)&S->pool[0], pool_walks * sizeof(gw_walk), "serve pool")) ||
        (rc = gw_alloc(ctx, (void**)&S->pool[1], pool_walks * sizeof(gw_walk), "serve pool (alt)")) ||
        (rc = gw_alloc(ctx, (void**)&S->arena, S->arena_slots * 8, "serve count arena")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_occ, S->arena_slots * 4, "serve occupied lists")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_occn, max_queries * 4ull, "serve occupied list lengths")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_slots, max_queries * sizeof(SlotInfo), "serve slots")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_remaining, max_queries * 4ull, "serve remaining")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_srccount, max_queries * 4ull, "serve source counts")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_cnt, 8 * 8, "serve counters")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_done, max_queries * sizeof(Done), "serve completions")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_region_n, max_queries * 8ull, "serve region counts")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_koff, max_queries * 8ull, "serve key regions")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_kend, max_queries * 8ull, "serve key region ends")) ||
        (rc = gw_alloc(ctx, (void**)&S->d_kcur, max_queries * 8ull, "serve key cursors"))) {
        gw_serve_close_state(ctx, S);
        return rc;
    }
    // the completion buffer, sized here so a step rarely grows it (an
    // allocation inside the serving loop would stall that step)
    S->pairs_cap = std::max<uint64_t>(1, std::min<uint64_t>(S->arena_slots, 1ull << 24));
    if ((rc = gw_alloc(ctx, (void**)&S->d_pairs, S->pairs_cap * 8, "serve compaction"))) {
        S->pairs_cap = 0;
        gw_serve_close_state(ctx, S);
        return rc;
    }
    {  // pinned staging of the per-step copies
        const uint64_t mq = max_queries;
        const uint64_t sizes[] = {mq * sizeof(Admit), mq * sizeof(Done), 8 * 8, mq * 4, mq * 8, mq * 4};
        uint64_t at[6], total = 0;
        for (int i = 0; i < 6; ++i) {
            at[i] = total;
            total += (sizes[i] + 63) / 64 * 64;
        }
        // coherent: kernels read the admissions and write the counters here,
        // and the host reads the kernels' writes after synchronising
        if (hipHostMalloc(&S->h_block, total, hipHostMallocCoherent) != hipSuccess) {
            S->h_block = nullptr;
            gw_serve_close_state(ctx, S);
            return gw_fail(ctx, GW_ENOMEM, "gw_serve_open: pinned staging buffer");
        }
        char* b = static_cast<char*>(S->h_block);
        S->h_adm = reinterpret_cast<Admit*>(b + at[0]);
        S->h_done = reinterpret_cast<Done*>(b + at[1]);
        S->h_cnt = reinterpret_cast<unsigned long long*>(b + at[2]);
        S->h_rem = reinterpret_cast<uint32_t*>(b + at[3]);
        S->h_kcur = reinterpret_cast<unsigned long long*>(b + at[4]);
        S->h_tn = reinterpret_cast<uint32_t*>(b + at[5]);
        char* db = nullptr;
        if (hipHostGetDevicePointer(reinterpret_cast<void**>(&db), S->h_block, 0) != hipSuccess || !db) {
            gw_serve_close_state(ctx, S);
            return gw_fail(ctx, GW_ENOMEM, "gw_serve_open: device address of the pinned staging buffer");
        }
        S->dh_adm = reinterpret_cast<Admit*>(db + at[0]);
        S->dh_cnt = reinterpret_cast<unsigned long long*>(db + at[2]);
        S->dh_rem = reinterpret_cast<uint32_t*>(db + at[3]);
        S->dh_kcur = reinterpret_cast<unsigned long long*>(db + at[4]);
        S->dh_tn = reinterpret_cast<uint32_t*>(db + at[5]);
    }
    if (ctx->host_graph && (rc = ooc_open(ctx, *S))) {
        gw_serve_close_state(ctx, S);
        return rc;
    }
    S->ext = !ctx->host_graph && prog->layout != GW_LAYOUT_CSR && gw_ensure_ext(ctx);
    if (prog->layout == GW_LAYOUT_EXT && !S->ext) {
        gw_serve_close_state(ctx, S);
        return gw_fail(ctx, GW_ENOMEM, "gw_serve_open: successor-header layout does not fit");
    }
    {  // workgroups of the serve-step kernel resident per CU (its LDS cache bounds them)
        int per_cu = 0;
        hipError_t oe = S->ooc   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_step_kernel<GW_RESTART, false, true>, 256, 0)
                      : S->ext ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_step_kernel<GW_RESTART, true, false>, 256, 0)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_step_kernel<GW_RESTART, false, false>, 256, 0);
        if (oe != hipSuccess || per_cu <= 0) per_cu = 4;
        S->grid_max = std::max(1u, ctx->cu_count) * static_cast<uint32_t>(std::min(per_cu, 8));
    }
    GW_HIP(ctx, hipMemsetAsync(S->arena, 0, S->arena_slots * 8, ctx->stream));
    GW_HIP(ctx, hipMemsetAsync(S->d_cnt, 0, 8 * 8, ctx->stream));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t i = 0; i < max_queries; ++i) S->free_slots.push_back(i);  // ascending: already a min-heap
    S->free_ranges.push_back({0, S->arena_slots});
    ctx->serve_state = S;
    return GW_OK;
}

int gw_serve_submit(gw_ctx* ctx, const gw_query* q, uint32_t n) {
    if (!ctx || (n && !q)) return gw_fail(ctx, GW_EINVAL, "gw_serve_submit: bad arguments");
    ServeState* sp = state_of(ctx);
    if (!sp) return gw_fail(ctx, GW_ESTATE, "gw_serve_submit: not open");
    ServeState& S = *sp;
    if (S.failed) return gw_fail(ctx, GW_ESTATE, "gw_serve_submit: the server failed; close it");
    std::set<uint64_t> batch;
    for (uint32_t i = 0; i < n; ++i) {
        if (q[i].source >= ctx->n || q[i].walks == 0 || q[i].walks > S.pool_cap)
            return gw_fail(ctx, GW_EINVAL, "gw_serve_submit: query " + std::to_string(q[i].qid) +
                                               " has a source out of range or a walk count outside [1, pool]");
        if (query_slots(q[i].walks, S.prog.L, ctx->n) > S.arena_slots)
            return gw_fail(ctx, GW_EINVAL, "gw_serve_submit: query " + std::to_string(q[i].qid) + " needs a count region of " +
                                               std::to_string(query_slots(q[i].walks, S.prog.L, ctx->n)) +
                                               " slots, more than the whole arena (" + std::to_string(S.arena_slots) + ")");
        if (S.qids.count(q[i].qid) || !batch.insert(q[i].qid).second)
            return gw_fail(ctx, GW_EINVAL, "gw_serve_submit: query id " + std::to_string(q[i].qid) +
                                               " is already queued, in flight or waiting to be read");
    }
    for (uint32_t i = 0; i < n; ++i) {
        S.queue.push_back(q[i]);
        S.qids.insert(q[i].qid);
    }
    return GW_OK;
}

int gw_serve_step(gw_ctx* ctx, gw_serve_stats* out) {
    if (!ctx) return GW_EINVAL;
    ServeState* sp = state_of(ctx);
    if (!sp) return gw_fail(ctx, GW_ESTATE, "gw_serve_step: not open");
    ServeState& S = *sp;
    if (S.failed) return gw_fail(ctx, GW_ESTATE, "gw_serve_step: the server failed; close it");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_serve_stats st{};
    // 0. over a host-resident graph: the page cache's table and loads
    if (S.ooc) {
        int rc = ooc_before_step(ctx, S, &st.h2d_bytes, &st.pages_loaded);
        if (rc) {
            S.failed = true;
            return rc;
        }
    }
    // 1. admit queued queries in arrival order while walk slots, a query slot
    //    and arena space for their count region are free
    const uint32_t L = S.prog.L;
    // with a top-k and hops >= L, every walk ends in the serve step that
    // starts it: the step completes what it admits, with one synchronisation
    const bool one_step = S.topk && S.hops >= L;
    std::vector<Admit> adm;
    while (!S.queue.empty()) {
        const gw_query& q = S.queue.front();
        if (S.live + q.walks > S.pool_cap || S.free_slots.empty()) break;
        const uint64_t size = region_slots(q.walks, L, ctx->n), total = query_slots(q.walks, L, ctx->n);
        uint64_t base = 0;
        if (total > S.arena_slots || !arena_alloc(S, total, base)) break;
        std::pop_heap(S.free_slots.begin(), S.free_slots.end(), std::greater<uint32_t>());
        const uint32_t slot = S.free_slots.back();
        S.free_slots.pop_back();
        const SlotInfo si{base, static_cast<uint32_t>(size - 1), q.source, static_cast<uint32_t>(q.qid),
                          static_cast<uint32_t>(q.qid >> 32), 0, 0, 0};  // srow, sdeg: written by the admit kernel
        const uint64_t koff = 2 * (base + size);
        adm.push_back(Admit{si, S.live, slot, q.walks, koff, koff + static_cast<uint64_t>(q.walks) * L});
        S.live += q.walks;
        S.active.push_back({q.qid, slot, q.walks, base, size, total, q.source});
        S.queue.pop_front();
    }
    if (!adm.empty()) {
        // the admit kernel reads the records from pinned memory; the previous
        // step's kernels are done with it (the step ended synchronised)
        std::copy(adm.begin(), adm.end(), S.h_adm);
        serve_admit_kernel<<<dim3(static_cast<uint32_t>(adm.size())), dim3(256), 0, ctx->stream>>>(
            S.dh_adm, S.pool[S.cur], S.d_slots, S.d_remaining, S.d_srccount, S.d_occn, S.ooc ? S.d_off : ctx->d_off,
            S.d_koff, S.d_kend, S.d_kcur);
        GW_HIP(ctx, hipGetLastError());
        st.admitted = static_cast<uint32_t>(adm.size());
    }
    // 2. one step over every live walk
    // the counters are zero: cleared at open, then by each step's export kernel
    GW_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if (S.live) {
        const Prog p{L, S.prog.nvertices, S.prog.stop_thr, static_cast<uint32_t>(S.prog.seed),
                     static_cast<uint32_t>(S.prog.seed >> 32), 0u, 0u, 0u};
        // one walk per lane while the pool fits the resident grid: no more
        // workgroups than the walks need (each wavefront that finds no work
        // still pays a cursor atomic), no more than stay resident at once
        const uint64_t need = (S.live + 255) / 256;
        const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(S.grid_max, need)));
        hipError_t e = S.prog.kind == GW_RESTART ? launch_serve<GW_RESTART>(ctx, S, p, S.live, grid)
                     : S.prog.kind == GW_STOP    ? launch_serve<GW_STOP>(ctx, S, p, S.live, grid)
                                                 : launch_serve<GW_JUMP>(ctx, S, p, S.live, grid);
        if (e != hipSuccess) return gw_hip_check(ctx, e, "serve step launch");
    }
    GW_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    const unsigned long long* cnt = S.h_cnt;
    const uint32_t* rem = S.h_rem;
    const unsigned long long* nkeys = S.h_kcur;
    // the remaining walks and key cursors of the slots in use only: slots are
    // handed out lowest first, so they sit below a small bound (the copies'
    // latency grows with their size)
    uint32_t hi = 0;
    for (const auto& a : S.active) hi = std::max(hi, a.slot + 1);
    const uint32_t np = S.ooc ? static_cast<uint32_t>(S.npages) : 0u;
    serve_export_kernel<<<dim3(std::max(1u, std::min(64u, (std::max(hi, np) + 255) / 256))), dim3(256), 0,
                          ctx->stream>>>(S.d_cnt, S.d_remaining, S.d_kcur, hi, S.dh_cnt, S.dh_rem, S.dh_kcur,
                                         S.d_page_hits, S.dh_hits, np);
    GW_HIP(ctx, hipGetLastError());
    // After the walks: the step's counters (overflow, steps, visits, walks
    // left) and, over a host graph, the page demand.
    auto after_walks = [&]() -> int {
        float ms = 0.f;
        GW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        if (cnt[3]) {  // the walks have moved on: the server cannot be stepped again consistently
            S.failed = true;
            return gw_fail(ctx, GW_ENOMEM, "gw_serve_step: a query's count region overflowed; close the server");
        }
        if (S.ooc) {
            for (uint64_t pg = 0; pg < S.npages; ++pg) {
                st.sampled_reads += S.hits[pg];
                if (S.page_slot[pg] >= 0) st.sampled_resident += S.hits[pg];
            }
            for (uint64_t pg = 0; pg < S.npages; ++pg) st.pages_resident += S.page_slot[pg] >= 0;
        }
        st.steps = cnt[1];
        st.visits = cnt[2];
        st.step_ms = ms;
        S.live = S.live ? cnt[0] : 0;
        S.cur ^= 1;
        return GW_OK;
    };
    // 3. completed queries: their keys are folded into their hash regions,
    //    whose entries (named by the occupied lists) are gathered back to
    //    back, with their slots cleared for reuse, then (with a top-k) the
    //    top-k of each is selected.  A query's pair count is only known on
    //    the device: its output range is sized by a bound on its key count,
    //    the key count itself (read back) or walks x L; the fold reads the
    //    true count on the device.
    std::vector<Done> dn;
    uint64_t total = 0;
    auto enqueue_completion = [&](const std::vector<ServeState::Active>& fin, bool bound_only) -> int {
        const uint32_t nd = static_cast<uint32_t>(fin.size());
        dn.assign(nd, Done{});
        total = 0;
        uint64_t maxkeys = 0, ncand = 0, maxc = 1;
        for (uint32_t i = 0; i < nd; ++i) {
            const auto& a = fin[i];
            // the cursor never passes the region (an overflow fails the step)
            const uint64_t cap = static_cast<uint64_t>(a.walks) * L;
            const uint64_t nk = bound_only ? cap : std::min<uint64_t>(nkeys[a.slot], cap);
            // top-k: the query's chunks' candidates follow the previous query's
            const uint64_t ch = (nk + 1 + kServeTopkChunk - 1) / kServeTopkChunk;
            dn[i] = Done{a.base, a.size, total, a.slot, a.source, 2 * (a.base + a.size), nk, ncand};
            total += nk;
            maxkeys = std::max<uint64_t>(maxkeys, nk);
            ncand += ch * S.topk;
            maxc = std::max(maxc, ch);
        }
        if (total > S.pairs_cap) {
            gw_free(ctx, S.d_pairs, S.pairs_cap * 8);
            S.d_pairs = nullptr;
            S.pairs_cap = 0;
            int rc = gw_alloc(ctx, (void**)&S.d_pairs, std::max<uint64_t>(total, 1 << 20) * 8, "serve compaction");
            if (rc) return rc;
            S.pairs_cap = std::max<uint64_t>(total, 1 << 20);
        }
        if (S.topk && ncand > S.cand_cap) {
            gw_free(ctx, S.d_cand, S.cand_cap * 8);
            S.d_cand = nullptr;
            S.cand_cap = 0;
            const uint64_t cap = std::max<uint64_t>(ncand, static_cast<uint64_t>(S.max_queries) * S.topk);
            int rc = gw_alloc(ctx, (void**)&S.d_cand, cap * 8, "serve top-k candidates");
            if (rc) return rc;
            S.cand_cap = cap;
        }
        std::copy(dn.begin(), dn.end(), S.h_done);
        GW_HIP(ctx, hipMemcpyAsync(S.d_done, S.h_done, nd * sizeof(Done), hipMemcpyHostToDevice, ctx->stream));
        if (maxkeys) {
            const uint32_t fx = static_cast<uint32_t>((maxkeys + kFoldKeys - 1) / kFoldKeys);
            serve_fold_kernel<<<dim3(fx, nd), dim3(512), 0, ctx->stream>>>(S.arena, S.d_occ, S.d_occn, S.d_done,
                                                                         S.d_kcur, S.cbits, S.d_cnt + 4);
            GW_HIP(ctx, hipGetLastError());
        }
        const uint32_t gx = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(64, (maxkeys + 255) / 256)));
        serve_gather_kernel<<<dim3(gx, nd), dim3(256), 0, ctx->stream>>>(S.arena, S.d_occ, S.d_occn, S.d_done,
                                                                         S.d_region_n, S.d_pairs);
        GW_HIP(ctx, hipGetLastError());
        if (S.topk) {
            // top-k per completed query on the device: k pairs per query cross PCIe
            serve_topk_chunk_kernel<<<dim3(static_cast<uint32_t>(maxc), nd), dim3(1024), 0, ctx->stream>>>(
                S.d_pairs, S.d_done, S.d_region_n, S.d_srccount, S.cbits, S.topk, S.d_cand);
            GW_HIP(ctx, hipGetLastError());
            // results, pair counts and the fold's overflow count land in pinned memory
            serve_topk_merge_kernel<<<dim3(nd), dim3(1024), 0, ctx->stream>>>(S.d_region_n, S.d_done, S.d_cand,
                                                                            S.cbits, S.topk, S.dh_top, S.dh_tn,
                                                                            S.d_cnt + 4, S.dh_cnt + 4);
            GW_HIP(ctx, hipGetLastError());
        }
        return GW_OK;
    };
    // The completed queries' top-k, after the synchronisation that follows
    // enqueue_completion: results kept, arena ranges and slots freed.
    auto take_topk = [&](const std::vector<ServeState::Active>& fin) -> int {
        if (S.h_cnt[4]) {  // the fold's overflow count, copied by the merge kernel
            S.failed = true;
            return gw_fail(ctx, GW_ENOMEM, "gw_serve_step: a query's count region overflowed; close the server");
        }
        const unsigned long long* top = S.h_top;
        const uint32_t* tn = S.h_tn;
        for (uint32_t i = 0; i < fin.size(); ++i) {
            const auto& a = fin[i];
            const auto* first = top + static_cast<uint64_t>(i) * S.topk;
            S.done[a.qid].assign(first, first + tn[i]);
            S.done_order.push_back(a.qid);
            arena_free(S, a.base, a.total);
            S.free_slots.push_back(a.slot);
            std::push_heap(S.free_slots.begin(), S.free_slots.end(), std::greater<uint32_t>());
        }
        st.completed = static_cast<uint32_t>(fin.size());
        return GW_OK;
    };
    int rc = GW_OK;
    if (one_step) {
        // Every walk ends within the serve step that starts it (hops >= L,
        // and a query is admitted whole), so this step completes exactly the
        // queries it admitted: their completion is enqueued behind the walks
        // and the step synchronises once.
        const std::vector<ServeState::Active> fin(S.active);
        if (!fin.empty() && (rc = enqueue_completion(fin, true))) return rc;
        GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if ((rc = after_walks())) return rc;
        for (const auto& a : fin)
            if (rem[a.slot] != 0) {
                S.failed = true;
                return gw_fail(ctx, GW_ESTATE, "gw_serve_step: a query's walks outlived their serve step");
            }
        if (!fin.empty() && (rc = take_topk(fin))) return rc;
        S.active.clear();
    } else {
        GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if ((rc = after_walks())) return rc;
        std::vector<ServeState::Active> still, fin;
        for (const auto& a : S.active) (rem[a.slot] != 0 ? still : fin).push_back(a);
        if (!fin.empty()) {
            if ((rc = enqueue_completion(fin, false))) return rc;
            const uint32_t nd = static_cast<uint32_t>(fin.size());
            if (S.topk) {
                GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
                if ((rc = take_topk(fin))) return rc;
            } else {
                unsigned long long fold_overflow = 0;
                std::vector<unsigned long long> per(nd);
                std::vector<unsigned long long> all(total);
                std::vector<uint32_t> srcc(S.max_queries);
                GW_HIP(ctx, hipMemcpyAsync(per.data(), S.d_region_n, nd * 8ull, hipMemcpyDeviceToHost, ctx->stream));
                if (total)
                    GW_HIP(ctx, hipMemcpyAsync(all.data(), S.d_pairs, total * 8, hipMemcpyDeviceToHost, ctx->stream));
                GW_HIP(ctx, hipMemcpyAsync(srcc.data(), S.d_srccount, S.max_queries * 4ull, hipMemcpyDeviceToHost,
                                           ctx->stream));
                GW_HIP(ctx, hipMemcpyAsync(&fold_overflow, S.d_cnt + 4, 8, hipMemcpyDeviceToHost, ctx->stream));
                GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
                if (fold_overflow) {
                    S.failed = true;
                    return gw_fail(ctx, GW_ENOMEM, "gw_serve_step: a query's count region overflowed; close the server");
                }
                for (uint32_t i = 0; i < nd; ++i) {
                    const auto& a = fin[i];
                    S.done[a.qid].assign(all.begin() + dn[i].out, all.begin() + dn[i].out + per[i]);
                    if (srcc[a.slot])
                        S.done[a.qid].push_back((static_cast<unsigned long long>(a.source) << S.cbits) | srcc[a.slot]);
                    S.done_order.push_back(a.qid);
                    arena_free(S, a.base, a.total);
                    S.free_slots.push_back(a.slot);
                    std::push_heap(S.free_slots.begin(), S.free_slots.end(), std::greater<uint32_t>());
                }
                st.completed = nd;
            }
        }
        S.active.swap(still);
    }
    st.live_walks = S.live;
    st.active_queries = S.active.size();
    st.queued = S.queue.size();
    if (out) *out = st;
    return GW_OK;
}

int gw_serve_poll(gw_ctx* ctx, uint64_t* qid, uint64_t* npairs) {
    if (!ctx || !qid || !npairs) return gw_fail(ctx, GW_EINVAL, "gw_serve_poll: bad arguments");
    ServeState* sp = state_of(ctx);
    if (!sp) return gw_fail(ctx, GW_ESTATE, "gw_serve_poll: not open");
    ServeState& S = *sp;
    // queries read by id through gw_serve_result before they were polled are skipped
    while (!S.done_order.empty() && !S.done.count(S.done_order.front())) S.done_order.pop_front();
    if (S.done_order.empty()) return 1;  // nothing completed
    const uint64_t id = S.done_order.front();
    S.done_order.pop_front();
    auto d = S.done.find(id);
    *qid = id;
    *npairs = d->second.size();
    return GW_OK;
}

int gw_serve_results(gw_ctx* ctx, uint32_t max_q, uint64_t* qid, uint64_t* npairs, uint32_t* vertex, uint32_t* count,
                     uint64_t cap, uint32_t* n_out) {
    if (!ctx || !n_out || (max_q && (!qid || !npairs)) || (cap && (!vertex || !count)))
        return gw_fail(ctx, GW_EINVAL, "gw_serve_results: bad arguments");
    ServeState* sp = state_of(ctx);
    if (!sp) return gw_fail(ctx, GW_ESTATE, "gw_serve_results: not open");
    ServeState& S = *sp;
    const unsigned long long cmask = (1ull << S.cbits) - 1;
    uint32_t nq = 0;
    uint64_t at = 0;
    while (nq < max_q && !S.done_order.empty()) {
        const uint64_t id = S.done_order.front();
        auto d = S.done.find(id);
        if (d == S.done.end()) {  // read already through gw_serve_result
            S.done_order.pop_front();
            continue;
        }
        const std::vector<unsigned long long>& pairs = d->second;
        if (at + pairs.size() > cap) {
            if (nq == 0) return gw_fail(ctx, GW_EINVAL, "gw_serve_results: the next query's pairs exceed cap");
            break;
        }
        for (size_t i = 0; i < pairs.size(); ++i) {
            vertex[at + i] = static_cast<uint32_t>(pairs[i] >> S.cbits);
            count[at + i] = static_cast<uint32_t>(pairs[i] & cmask);
        }
        qid[nq] = id;
        npairs[nq] = pairs.size();
        at += pairs.size();
        ++nq;
        S.done.erase(d);
        S.qids.erase(id);
        S.done_order.pop_front();
    }
    *n_out = nq;
    return GW_OK;
}

int gw_serve_result(gw_ctx* ctx, uint64_t qid, uint32_t* vertex, uint32_t* count, uint64_t cap) {
    if (!ctx) return GW_EINVAL;
    ServeState* sp = state_of(ctx);
    if (!sp) return gw_fail(ctx, GW_ESTATE, "gw_serve_result: not open");
    ServeState& S = *sp;
    auto d = S.done.find(qid);
    if (d == S.done.end()) return gw_fail(ctx, GW_EINVAL, "gw_serve_result: no completed query with that id");
    std::vector<unsigned long long>& pairs = d->second;
    if (cap < pairs.size() || (!pairs.empty() && (!vertex || !count)))
        return gw_fail(ctx, GW_EINVAL, "gw_serve_result: output arrays too small");
    const unsigned long long cmask = (1ull << S.cbits) - 1;
    for (size_t i = 0; i < pairs.size(); ++i) {
        vertex[i] = static_cast<uint32_t>(pairs[i] >> S.cbits);
        count[i] = static_cast<uint32_t>(pairs[i] & cmask);
    }
    S.done.erase(d);
    S.qids.erase(qid);
    return GW_OK;
}

int gw_serve_topk(gw_ctx* ctx, uint32_t k) {
    if (!ctx) return GW_EINVAL;
    ServeState* sp = state_of(ctx);
    if (!sp) return gw_fail(ctx, GW_ESTATE, "gw_serve_topk: not open");
    if (k > kTopkMax) return gw_fail(ctx, GW_EINVAL, "gw_serve_topk: k must be at most 1024");
    ServeState& S = *sp;
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    gw_free(ctx, S.d_cand, S.cand_cap * 8);
    S.d_cand = nullptr;
    if (S.h_top) (void)hipHostFree(S.h_top);
    S.h_top = S.dh_top = nullptr;
    S.cand_cap = 0;
    S.topk = 0;
    if (k == 0) return GW_OK;
    int rc;
    // coherent: the merge kernel writes the results here
    if (hipHostMalloc(reinterpret_cast<void**>(&S.h_top), static_cast<uint64_t>(S.max_queries) * k * 8,
                      hipHostMallocCoherent) != hipSuccess) {
        S.h_top = nullptr;
        return gw_fail(ctx, GW_ENOMEM, "gw_serve_topk: pinned result buffer");
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&S.dh_top), S.h_top, 0) != hipSuccess || !S.dh_top) {
        (void)hipHostFree(S.h_top);
        S.h_top = S.dh_top = nullptr;
        return gw_fail(ctx, GW_ENOMEM, "gw_serve_topk: device address of the pinned result buffer");
    }
    // candidates for a step whose completing queries average up to 8 chunks
    // (64 K pairs each); a step that needs more grows the buffer
    const uint64_t cand = static_cast<uint64_t>(S.max_queries) * 8 * k;
    if ((rc = gw_alloc(ctx, (void**)&S.d_cand, cand * 8, "serve top-k candidates"))) return rc;
    S.cand_cap = cand;
    S.topk = k;
    return GW_OK;
}

int gw_serve_cache(gw_ctx* ctx, uint32_t page_log2, uint32_t max_loads) {
    if (!ctx || (page_log2 != 0 && (page_log2 < 10 || page_log2 > 26)) || max_loads > 64)
        return gw_fail(ctx, GW_EINVAL, "gw_serve_cache: page_log2 in [10, 26] (or 0), max_loads at most 64");
    if (state_of(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_serve_cache: call it before gw_serve_open");
    ctx->serve_page_log2 = page_log2;
    ctx->serve_max_loads = max_loads;
    return GW_OK;
}

int gw_serve_close(gw_ctx* ctx) {
    if (!ctx) return GW_EINVAL;
    ServeState* S = state_of(ctx);
    if (!S) return GW_OK;
    ctx->serve_state = nullptr;
    gw_serve_close_state(ctx, S);
    return GW_OK;
}

}  // extern "C"

static void gw_serve_close_state(gw_ctx* ctx, gw::ServeState* S) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (S->ooc) gw::ooc_close(ctx, *S);
    gw_free(ctx, S->pool[0], S->pool_cap * sizeof(gw_walk));
    gw_free(ctx, S->pool[1], S->pool_cap * sizeof(gw_walk));
    gw_free(ctx, S->arena, S->arena_slots * 8);
    gw_free(ctx, S->d_occ, S->arena_slots * 4);
    gw_free(ctx, S->d_occn, S->max_queries * 4ull);
    gw_free(ctx, S->d_slots, S->max_queries * sizeof(SlotInfo));
    gw_free(ctx, S->d_remaining, S->max_queries * 4ull);
    gw_free(ctx, S->d_srccount, S->max_queries * 4ull);
    gw_free(ctx, S->d_cnt, 64);
    gw_free(ctx, S->d_pairs, S->pairs_cap * 8);
    gw_free(ctx, S->d_done, S->max_queries * sizeof(Done));
    gw_free(ctx, S->d_region_n, S->max_queries * 8ull);
    if (S->h_block) (void)hipHostFree(S->h_block);
    if (S->h_top) (void)hipHostFree(S->h_top);
    gw_free(ctx, S->d_koff, S->max_queries * 8ull);
    gw_free(ctx, S->d_kend, S->max_queries * 8ull);
    gw_free(ctx, S->d_kcur, S->max_queries * 8ull);
    gw_free(ctx, S->d_cand, S->cand_cap * 8);
    delete S;
}
