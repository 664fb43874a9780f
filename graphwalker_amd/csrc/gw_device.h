// gw_device.h -- device-side building blocks of the walk engine (gfx950).
//
//  * Philox4x32-10 (Salmon et al., SC'11): counter-based, so a walk's random
//    stream is a pure function of (seed, walk id, hop) and trajectories do not
//    depend on the launch shape or the order walks are scheduled in.
//  * The per-step semantics of the reference's updateByWalk variants
//    (src/walks/randomwalkwith*.hpp), written as one templated advance() whose
//    else-branch is selected at compile time.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gwhip.h"

namespace gw {

constexpr uint32_t kTagWalk = 0x57414C4Bu;  // 'WALK'
constexpr uint32_t kTagRmat = 0x524D4154u;  // 'RMAT'
constexpr uint32_t kTagSrcs = 0x53524353u;  // 'SRCS'
constexpr uint64_t kHashInit = 0xCBF29CE484222325ull;
constexpr uint64_t kHashPrime = 0x100000001B3ull;

struct u32x4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ uint32_t umulhi32(uint32_t a, uint32_t b) {
    return static_cast<uint32_t>((static_cast<uint64_t>(a) * b) >> 32);
}

__host__ __device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                         uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32 x 32 -> 64-bit product per multiplier (v_mad_u64_u32 on
        // gfx950) instead of separate high and low multiplies
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = static_cast<uint32_t>(p1); c2 = n2; c3 = static_cast<uint32_t>(p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

__device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) { return __umulhi(a, b); }

// Lane `lane`'s value of x, for a wave-uniform lane index (v_readlane into a
// scalar register: no LDS crossbar round trip as a shuffle would take).
__device__ __forceinline__ uint32_t lane_value(uint32_t x, uint32_t lane) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), static_cast<int>(lane)));
}
__device__ __forceinline__ unsigned long long lane_value(unsigned long long x, uint32_t lane) {
    const uint32_t lo = lane_value(static_cast<uint32_t>(x), lane);
    const uint32_t hi = lane_value(static_cast<uint32_t>(x >> 32), lane);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// Program parameters as the kernels see them (kind is a template argument).
struct Prog {
    uint32_t L;
    uint32_t nvertices;
    uint64_t stop_thr;
    uint32_t k0, k1;   // Philox key = seed
    uint32_t stream;
    // Per-walk outputs (final record, summary, event log) of walk id w are at
    // index w - id_base: a window of a host walk pool writes window-sized
    // arrays.  The random stream stays keyed by the walk id itself.
    uint32_t id_base;
    uint32_t chunk;  // walks a wavefront claims per atomic (persistent and blocked schedules)
};

// Successor-header edge (16 bytes): the column entry plus the destination's
// degree and row start, so one random read per step yields everything the
// next step needs (the CSR layout needs two dependent reads: offsets, column).
struct __attribute__((aligned(16))) Edge {
    uint32_t dst;
    uint32_t deg;
    uint64_t row;
};

struct Graph {
    const uint64_t* __restrict__ off;
    const uint32_t* __restrict__ col;
    const Edge* __restrict__ ext;  // null unless the successor-header layout is built
    uint32_t n;
};

enum : int { kMoved = 0, kFinished = 1 };

// Per-walk event log: visit k of walk `id` goes to slot id * cap + k as
// (hop << 32 | vertex); visits beyond `cap` are counted but not stored.
struct EventLog {
    uint64_t* __restrict__ ev;
    uint32_t cap;
    uint32_t keep_last;  // visits past cap overwrite the last slot (cap >= 1)
};

// Device-wide open-addressing hash of (source field, vertex) -> visit count,
// key = (source << vbits) | vertex.  Packed (keys of at most 44 bits): a slot
// holds (key << cbits) | count, 0 marks an empty slot (every stored entry has
// count >= 1).  Wide: a slot is two words {key + 1, count}, 0 marks empty.
// Linear probing from a multiplicative hash; a full table raises `overflow`.
struct SourceTable {
    unsigned long long* __restrict__ tab;
    unsigned long long* __restrict__ overflow;
    uint64_t mask;
    uint32_t shift;  // 64 - log2(slots)
    uint32_t vbits;
    uint32_t cbits;
    uint32_t wide;
};

// Adds `inc` visits to (source, v).
__device__ __forceinline__ void table_add(const SourceTable& t, uint32_t source, uint32_t v, uint32_t inc = 1) {
    const uint64_t key = (static_cast<uint64_t>(source) << t.vbits) | v;
    uint64_t h = (key * 0x9E3779B97F4A7C15ull) >> t.shift;
    if (t.wide) {
        const unsigned long long tag = key + 1;
        for (uint64_t probe = 0; probe <= t.mask; ++probe) {
            const unsigned long long old = atomicCAS(&t.tab[2 * h], 0ull, tag);
            if (old == 0ull || old == tag) {
                atomicAdd(&t.tab[2 * h + 1], static_cast<unsigned long long>(inc));
                return;
            }
            h = (h + 1) & t.mask;
        }
        atomicAdd(t.overflow, static_cast<unsigned long long>(inc));
        return;
    }
    const unsigned long long entry = (static_cast<unsigned long long>(key) << t.cbits) | inc;
    for (uint64_t probe = 0; probe <= t.mask; ++probe) {
        const unsigned long long old = atomicCAS(&t.tab[h], 0ull, entry);
        if (old == 0ull) return;
        if ((old >> t.cbits) == key) {
            atomicAdd(&t.tab[h], static_cast<unsigned long long>(inc));
            return;
        }
        h = (h + 1) & t.mask;
    }
    atomicAdd(t.overflow, static_cast<unsigned long long>(inc));
}

// Append-only log of (source, vertex) keys for sort-reduce aggregation.  Each
// wavefront owns a slab of `slab` consecutive slots (claimed with one global
// atomic) and fills it with the keys of its visiting lanes, ranked by ballot;
// keys that do not fit spill into the next slab.  At kernel exit the open
// slab's tail is padded with `sentinel`, which sorts after every key.
struct KeyLog {
    unsigned long long* __restrict__ keys;
    unsigned long long* __restrict__ cursor;
    unsigned long long* __restrict__ overflow;
    uint64_t cap;
    uint64_t sentinel;
    uint32_t vbits;
    uint32_t slab;
    // Direct regions (no log): visits of source field f go straight to
    // region_out[region_off[f] + ...), claimed through region_n[f], up to
    // region_end[f] (region_off + 1 where the regions are back to back).
    const uint64_t* __restrict__ region_off;
    const uint64_t* __restrict__ region_end;
    unsigned long long* __restrict__ region_n;
    uint32_t* __restrict__ region_out;
    unsigned long long* __restrict__ region_overflow;
    uint32_t nf;
    // a designated vertex per field (the first walk's start) whose visits are
    // counted in field_count[f] instead of the region: every walk of a
    // single-source field opens with one
    const uint32_t* __restrict__ field_vertex;
    uint32_t* __restrict__ field_count;
};

// Called by the lanes that visit; `slab` = this wavefront's {next, end} in LDS.
__device__ __forceinline__ void keylog_add(const KeyLog& k, uint64_t* slab, uint32_t source, uint32_t v) {
    const uint64_t key = (static_cast<uint64_t>(source) << k.vbits) | v;
    const uint64_t mask = __ballot(1);
    const uint32_t cnt = __popcll(mask);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
    const uint32_t leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    const uint64_t next = slab[0], end = slab[1];
    const uint32_t room = end - next < cnt ? static_cast<uint32_t>(end - next) : cnt;
    if (room == cnt) {
        k.keys[next + rank] = key;
        if (rank == 0) slab[0] = next + cnt;
        return;
    }
    unsigned long long b = 0;
    if (rank == 0) b = atomicAdd(k.cursor, static_cast<unsigned long long>(k.slab));
    b = lane_value(b, leader);
    const bool ok = b + k.slab <= k.cap;
    if (rank < room) k.keys[next + rank] = key;
    else if (ok) k.keys[b + (rank - room)] = key;
    if (rank == 0) {
        if (!ok) atomicAdd(k.overflow, 1ull);
        slab[0] = ok ? b + (cnt - room) : 0;
        slab[1] = ok ? b + k.slab : 0;
    }
}

// Direct regions, staged per wavefront in LDS: two buffers of kRegionBuf
// vertex ids, each holding visits of one source field.  A full buffer (or
// one needed for a third field) is flushed to its field's region with one
// cursor atomic for the whole buffer and coalesced stores, so the step
// kernel waits on no atomic per visit and leaves no holes in a region.
constexpr uint32_t kRegionBuf = 1024;
constexpr uint32_t kNoField = 0xFFFFFFFFu;
struct alignas(16) RegionBuf {
    uint32_t st[4];  // {field, count} of buffer 0, then of buffer 1: one 16-B LDS read
    uint32_t key[2][kRegionBuf];
};

// One lane per wavefront, at kernel start.
__device__ __forceinline__ void region_buf_init(RegionBuf* rb) {
    rb->st[0] = rb->st[2] = kNoField;
    rb->st[1] = rb->st[3] = 0;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Writes the n staged ids of field f to its region; the lanes in `active`
// take part.  Ids past the region's end (impossible while regions hold every
// visit a walk can make) count in region_overflow.  The caller records the
// buffer's new state.
__device__ __forceinline__ void region_flush(const KeyLog& k, const uint32_t* keys, uint32_t f, uint32_t n,
                                             uint64_t active) {
    if (n == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll(static_cast<unsigned long long>(active)) - 1;
    unsigned long long at = 0;
    if (lane == leader) at = atomicAdd(&k.region_n[f], static_cast<unsigned long long>(n));
    at = lane_value(at, leader);
    const uint64_t base = k.region_off[f] + at, hi = k.region_end[f];
    const uint32_t m = __popcll(active);
    for (uint32_t i = lanes_below(active); i < n; i += m)
        if (base + i < hi) k.region_out[base + i] = keys[i];
    if (lane == leader && base + n > hi)
        atomicAdd(k.region_overflow, static_cast<unsigned long long>(base + n - (base > hi ? base : hi)));
    __builtin_amdgcn_wave_barrier();  // every lane has read the ids before they are overwritten
}

// Called by the lanes that visit: each stages its vertex id in the buffer of
// its own source field.  The visiting lanes are grouped by field with a
// ballot loop (a wavefront's walks come from one stretch of the walk array,
// so they name one or two fields).  A field >= nf counts in region_overflow.
__device__ __forceinline__ void region_add(const KeyLog& k, RegionBuf* rb, uint32_t source, uint32_t v,
                                           uint32_t field_v) {
    const uint64_t active = __ballot(1);
    const uint32_t lane = __lane_id();
    // visits of the field's designated vertex: one count per (wavefront, field)
    const bool fvis = v == field_v;
    uint64_t fdo = __ballot(fvis);
    while (fdo) {  // wave-uniform
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(fdo)) - 1;
        const uint32_t ls = lane_value(source, leader);
        const uint64_t grp = __ballot(fvis && source == ls);
        if (lane == leader && ls < k.nf) atomicAdd(&k.field_count[ls], static_cast<uint32_t>(__popcll(grp)));
        fdo &= ~grp;
    }
    uint64_t todo = __ballot(!fvis);
    while (todo) {  // wave-uniform
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
        const uint32_t ls = lane_value(source, leader);
        const bool in = !fvis && source == ls;
        const uint64_t grp = __ballot(in);
        todo &= ~grp;
        const uint32_t g = __popcll(grp);
        if (ls >= k.nf) {
            if (lane == leader) atomicAdd(k.region_overflow, static_cast<unsigned long long>(g));
            continue;
        }
        const uint4 st = *reinterpret_cast<const uint4*>(rb->st);
        uint32_t b, c;
        if (st.x == ls) {
            b = 0;
            c = st.y;
        } else if (st.z == ls) {
            b = 1;
            c = st.w;
        } else {  // take an empty buffer, else flush the fuller one
            b = st.y == 0 ? 0u : st.w == 0 ? 1u : st.y >= st.w ? 0u : 1u;
            region_flush(k, rb->key[b], b ? st.z : st.x, b ? st.w : st.y, active);
            c = 0;
        }
        if (c + g > kRegionBuf) {
            region_flush(k, rb->key[b], ls, c, active);
            c = 0;
        }
        if (in) rb->key[b][c + lanes_below(grp)] = v;
        __builtin_amdgcn_wave_barrier();
        if (lane == leader) *reinterpret_cast<uint2*>(&rb->st[2 * b]) = make_uint2(ls, c + g);
        __builtin_amdgcn_wave_barrier();
    }
}

// The same direct regions for a caller that records in wave-uniform control
// flow (every lane of the wavefront present, `want` marking the lanes with a
// visit): the two buffers' {field, count} state lives in wave-uniform
// registers instead of LDS, and group leaders broadcast with readlane.  Visits
// of a buffered field's designated vertex are only counted, in d0 / d1, and
// reach field_count when the field leaves its buffer (or at kernel exit):
// nearly every step of a short-walk batch starts walks at their source.  A
// buffer whose field has only such counts (c == 0) counts as empty: giving
// it to another field evicts just the count.
struct RegionState {
    uint32_t f0 = kNoField, c0 = 0, f1 = kNoField, c1 = 0, d0 = 0, d1 = 0;
};

__device__ __forceinline__ void region_flush_all(const KeyLog& k, const uint32_t* keys, uint32_t f, uint32_t n) {
    if (n == 0) return;
    const uint32_t lane = __lane_id();
    unsigned long long at = 0;
    if (lane == 0) at = atomicAdd(&k.region_n[f], static_cast<unsigned long long>(n));
    at = lane_value(at, 0u);
    const uint64_t base = k.region_off[f] + at, hi = k.region_end[f];
    for (uint32_t i = lane; i < n; i += 64)
        if (base + i < hi) k.region_out[base + i] = keys[i];
    if (lane == 0 && base + n > hi)
        atomicAdd(k.region_overflow, static_cast<unsigned long long>(base + n - (base > hi ? base : hi)));
    __builtin_amdgcn_wave_barrier();  // every lane has read the ids before they are overwritten
}

// A buffered field leaves its buffer: its staged ids to the region, its
// designated-vertex count to field_count.
__device__ __forceinline__ void region_evict(const KeyLog& k, const uint32_t* keys, uint32_t f, uint32_t n,
                                             uint32_t d) {
    region_flush_all(k, keys, f, n);
    if (d && __lane_id() == 0) atomicAdd(&k.field_count[f], d);
}

// The state is wave-uniform by construction; readfirstlane tells the compiler
// so, which keeps it in scalar registers and its branches scalar (otherwise a
// divergent atomic's join turns the whole decision chain into exec-masked code).
__device__ __forceinline__ uint32_t uniform(uint32_t x) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
}

template <bool kCountsTakeBuffers = false>
__device__ __forceinline__ void region_add_uniform(const KeyLog& k, RegionBuf* rb, RegionState& st, bool want,
                                                   uint32_t source, uint32_t v, uint32_t field_v) {
    const uint32_t lane = __lane_id();
    st.f0 = uniform(st.f0);
    st.c0 = uniform(st.c0);
    st.f1 = uniform(st.f1);
    st.c1 = uniform(st.c1);
    st.d0 = uniform(st.d0);
    st.d1 = uniform(st.d1);
    // visits of the field's designated vertex: counted in the buffer state
    // while their field holds a buffer -- with kCountsTakeBuffers, a field
    // without one takes a buffer that holds no staged ids (evicting just its
    // count) -- else one atomic per (wavefront, field).  Serving takes
    // buffers: a query whose walks only ever revisit that vertex (restart
    // walks from a vertex with no out-edges) would otherwise add to one
    // address from every wavefront at every hop.  Batched restart walks do
    // the same; batched walks of the other kinds visit the designated vertex
    // once, at their start, and keep the plain rule.
    const bool fvis = want && v == field_v;
    uint64_t fdo = __ballot(fvis);
    while (fdo) {
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(fdo)) - 1;
        const uint32_t ls = lane_value(source, leader);
        const uint64_t grp = __ballot(fvis && source == ls);
        const uint32_t d = __popcll(grp);
        bool in0 = st.f0 == ls, in1 = st.f1 == ls;
        if (kCountsTakeBuffers && !in0 && !in1 && ls < k.nf) {  // wavefront-uniform
            if (st.c0 == 0) {
                if (st.d0 && lane == 0) atomicAdd(&k.field_count[st.f0], st.d0);
                st.f0 = ls;
                st.d0 = 0;
                in0 = true;
            } else if (st.c1 == 0) {
                if (st.d1 && lane == 0) atomicAdd(&k.field_count[st.f1], st.d1);
                st.f1 = ls;
                st.d1 = 0;
                in1 = true;
            }
        }
        st.d0 += in0 ? d : 0u;
        st.d1 += in1 ? d : 0u;
        if (!in0 && !in1 && ls < k.nf && lane == leader) atomicAdd(&k.field_count[ls], d);
        fdo &= ~grp;
        if (kCountsTakeBuffers) {
            st.f0 = uniform(st.f0);
            st.d0 = uniform(st.d0);
            st.f1 = uniform(st.f1);
            st.d1 = uniform(st.d1);
        }
    }
    const bool rest = want && !fvis;
    uint64_t todo = __ballot(rest);
    while (todo) {
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
        const uint32_t ls = lane_value(source, leader);
        const bool in = rest && source == ls;
        const uint64_t grp = __ballot(in);
        todo &= ~grp;
        const uint32_t g = __popcll(grp);
        if (ls >= k.nf) {
            if (lane == leader) atomicAdd(k.region_overflow, static_cast<unsigned long long>(g));
        } else {
            // the field's buffer: its own, else an empty one, else the fuller
            // one evicted; the state changes by selects, not branches
            const bool in0 = st.f0 == ls, in1 = st.f1 == ls, take = !in0 && !in1;
            const bool second = in1 || (take && st.c0 != 0 && (st.c1 == 0 || st.c1 > st.c0));
            if (take) {
                if (second) region_evict(k, rb->key[1], st.f1, st.c1, st.d1);
                else region_evict(k, rb->key[0], st.f0, st.c0, st.d0);
            }
            const bool new0 = take && !second, new1 = take && second;
            st.f0 = new0 ? ls : st.f0;
            st.c0 = new0 ? 0u : st.c0;
            st.d0 = new0 ? 0u : st.d0;
            st.f1 = new1 ? ls : st.f1;
            st.c1 = new1 ? 0u : st.c1;
            st.d1 = new1 ? 0u : st.d1;
            uint32_t* keys = rb->key[second ? 1 : 0];
            uint32_t c = second ? st.c1 : st.c0;
            const bool full = c + g > kRegionBuf;
            if (full) region_flush_all(k, keys, ls, c);
            c = full ? 0u : c;
            if (in) keys[c + lanes_below(grp)] = v;
            __builtin_amdgcn_wave_barrier();
            st.c0 = second ? st.c0 : c + g;
            st.c1 = second ? c + g : st.c1;
        }
        st.f0 = uniform(st.f0);
        st.c0 = uniform(st.c0);
        st.f1 = uniform(st.f1);
        st.c1 = uniform(st.c1);
        st.d0 = uniform(st.d0);
        st.d1 = uniform(st.d1);
    }
}

// At kernel exit (all lanes of the wavefront present): flush both buffers.
__device__ __forceinline__ void region_close(const KeyLog& k, RegionBuf* rb) {
    const uint64_t all = __ballot(1);
    const uint4 st = *reinterpret_cast<const uint4*>(rb->st);
    region_flush(k, rb->key[0], st.x, st.y, all);
    region_flush(k, rb->key[1], st.z, st.w, all);
}

// At kernel exit (all lanes of the wavefront present): pad the open slab.
__device__ __forceinline__ void keylog_close(const KeyLog& k, uint64_t* slab, uint32_t lane) {
    for (uint64_t i = slab[0] + lane; i < slab[1]; i += 64) k.keys[i] = k.sentinel;
}

// Recorder hooks, compile-time selected: the device replacements for the
// per-step updateInfo callback (randomwalk.hpp:32).
template <bool kVertexCount, bool kSummary, bool kEvents = false, bool kSourceCount = false, bool kKeyLog = false,
          bool kRegion = false, bool kDeferRegion = false, bool kCountsTakeBuffers = false>
struct Recorder {
    uint32_t* __restrict__ counts;
    uint32_t visits;
    uint32_t steps;
    uint64_t hash;
    EventLog log;
    uint32_t id;
    SourceTable table;
    uint32_t source;
    KeyLog klog;
    uint64_t* slab;
    RegionBuf* rbuf;
    uint32_t field_v = 0xFFFFFFFFu;  // the current walk's field's designated vertex (direct regions)
    // kDeferRegion: the step's visit, handed to the regions by commit() in
    // wave-uniform control flow, and the regions' buffer state
    uint32_t pend_v = 0;
    bool pend = false;
    RegionState rst;
    __device__ __forceinline__ Recorder(uint32_t* c, const EventLog& l, const SourceTable& t, const KeyLog& kl,
                                        uint64_t* sl, RegionBuf* rb = nullptr)
        : counts(c), visits(0), steps(0), hash(kHashInit), log(l), id(0), table(t), source(0), klog(kl), slab(sl),
          rbuf(rb) {}
    __device__ __forceinline__ void visit(uint32_t v, uint32_t hop) {
        if (kEvents && (visits < log.cap || log.keep_last))
            log.ev[static_cast<uint64_t>(id) * log.cap + (visits < log.cap ? visits : log.cap - 1)] =
                (static_cast<uint64_t>(hop) << 32) | v;
        ++visits;
        if (kVertexCount || kSourceCount) {
            // A walk standing on its own source field's vertex -- for restart
            // programs the restart target, a sixth or more of all visits --
            // would make every such lane of the wavefront hit one address.
            // Those lanes combine: one atomic per (wavefront, source).
            const bool at_src = v == source;
            uint64_t todo = __ballot(at_src);
            while (todo) {
                const uint32_t leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
                const uint32_t ls = lane_value(source, leader);
                const uint64_t grp = __ballot(at_src && source == ls);
                if (__lane_id() == leader) {
                    const uint32_t c = static_cast<uint32_t>(__popcll(grp));
                    if (kVertexCount) atomicAdd(&counts[ls], c);
                    if (kSourceCount) table_add(table, ls, ls, c);
                }
                todo &= ~grp;
            }
            if (!at_src) {
                if (kVertexCount) atomicAdd(&counts[v], 1u);
                if (kSourceCount) table_add(table, source, v);
            }
        }
        if (kSummary) hash = (hash ^ ((static_cast<uint64_t>(v) << 32) | hop)) * kHashPrime;
        if (kKeyLog) keylog_add(klog, slab, source, v);
        if (kRegion && kDeferRegion) {
            pend_v = v;
            pend = true;
        } else if (kRegion) {
            region_add(klog, rbuf, source, v, field_v);
        }
    }
    // kDeferRegion: every lane of the wavefront calls it once per step, after
    // the step (each step makes at most one visit per lane).
    __device__ __forceinline__ void commit() {
        if (kRegion && kDeferRegion) {
            region_add_uniform<kCountsTakeBuffers>(klog, rbuf, rst, pend, source, pend_v, field_v);
            pend = false;
        }
    }
    // At kernel exit, every lane present.
    __device__ __forceinline__ void close_regions() {
        if (kRegion && kDeferRegion) {
            region_evict(klog, rbuf->key[0], rst.f0, rst.c0, rst.d0);
            region_evict(klog, rbuf->key[1], rst.f1, rst.c1, rst.d1);
        } else if (kRegion) {
            region_close(klog, rbuf);
        }
    }
};

// Advance one walk while its vertex lies in [lo, hi).  Returns kFinished when
// the walk ends, kMoved when it left [lo, hi) alive.  Else-branch per kind:
//   GW_JUMP          teleport to mulhi32(r3, N)        randomwalkwithjump.hpp:49-51
//   GW_STOP/GW_PROB  die                                randomwalkwithstop.hpp:39-41
//   GW_RESTART_JOINT restart at source; end when hop % L == L-1
//                                                       randomwalkwithrestartwithjoint.hpp:47-53
//   GW_RESTART       restart at source, cap L           randomwalkwithrestart.hpp:33-63
// The loop guard (vertex in block, hop < L) is the while-condition of the
// reference loops (randomwalkwithjump.hpp:42).
template <int KIND, class Rec>
__device__ __forceinline__ int advance(const Prog& p, const Graph& g, uint32_t id, uint32_t source,
                                       uint32_t& v_io, uint32_t& hop_io, uint32_t lo, uint32_t hi,
                                       Rec& rec) {
    constexpr bool capped = (KIND == GW_JUMP || KIND == GW_STOP || KIND == GW_RESTART);
    uint32_t v = v_io, hop = hop_io;
    while (v - lo < hi - lo) {  // lo <= v < hi
        if (capped && hop >= p.L) { v_io = v; hop_io = hop; return kFinished; }
        rec.visit(v, hop);
        const u32x4 r = philox4x32_10(id, hop, kTagWalk, p.stream, p.k0, p.k1);
        const uint64_t b = g.off[v];
        const uint64_t deg = g.off[v + 1] - b;
        if (deg > 0 && static_cast<uint64_t>(r.x) >= p.stop_thr) {
            const uint64_t x = static_cast<uint64_t>(r.y) | (static_cast<uint64_t>(r.z) << 32);
            v = g.col[b + mulhi64(x, deg)];
        } else {
            if (KIND == GW_JUMP) {
                v = mulhi32(r.w, p.nvertices);
            } else if (KIND == GW_STOP || KIND == GW_PROB) {
                v_io = v; hop_io = hop;
                return kFinished;
            } else {
                v = source;
            }
        }
        ++hop;
        ++rec.steps;
        if (KIND == GW_RESTART_JOINT && hop % p.L == p.L - 1) { v_io = v; hop_io = hop; return kFinished; }
    }
    v_io = v;
    hop_io = hop;
    if (capped && hop >= p.L) return kFinished;
    return kMoved;
}

// Where a generated walk starts (walks placed per source: walk i belongs to
// source entry i / wps): its vertex, source field, the field's designated
// vertex, the vertex's row and degree, and the row and degree of the field as
// a vertex (the restart target of the restart programs; zero when the field
// is not a vertex).  48 bytes, read by a wavefront with scalar loads.
struct alignas(16) SourceStart {
    uint32_t vertex, field, fvertex, deg;
    uint64_t row;
    uint64_t srow;
    uint32_t sdeg, pad0;
    uint64_t pad1;
};
static_assert(sizeof(SourceStart) == 48, "SourceStart is 48 bytes");

struct WalkGen {
    const SourceStart* __restrict__ tab;  // null: the kernel reads walk records
    uint32_t wps;                          // walks per source entry
    uint32_t nsrc;                         // entries
};

// Walk state carried in registers.  With the successor-header layout the
// current vertex's row start and degree ride along, fetched with the edge
// that led to the vertex.
struct State {
    uint32_t id, source, v, hop;
    uint64_t row;
    uint32_t deg;
    u32x4 r;  // the Philox draw of (id, hop): set by start_walk, advanced by step_once
    // restart programs on the successor-header layout: the restart target's
    // row and degree, fetched once per walk (set_source_row) -- a restart
    // then needs no load, and in a wavefront of restart walks some lane
    // restarts on nearly every hop
    uint64_t srow;
    uint32_t sdeg;
};

// A walk picked up by a kernel: its draw for the current hop.
__device__ __forceinline__ void start_walk(const Prog& p, State& s) {
    s.r = philox4x32_10(s.id, s.hop, kTagWalk, p.stream, p.k0, p.k1);
}

template <bool EXT>
__device__ __forceinline__ void load_row(const Graph& g, State& s) {
    if (EXT) {
        const uint64_t b = g.off[s.v];
        s.row = b;
        s.deg = static_cast<uint32_t>(g.off[s.v + 1] - b);
    }
}

template <int KIND>
constexpr bool restarts() { return KIND == GW_RESTART || KIND == GW_RESTART_JOINT; }

// At a walk's start: the restart target's row and degree (restart programs
// on the successor-header layout; s.source must be a vertex of g).
template <int KIND, bool EXT>
__device__ __forceinline__ void set_source_row(const Graph& g, State& s) {
    if (EXT && restarts<KIND>()) {
        const uint64_t b = g.off[s.source];
        s.srow = b;
        s.sdeg = static_cast<uint32_t>(g.off[s.source + 1] - b);
    }
}

// Where the CSR layout's column of edge e is read from: the resident column
// array (PlainCols), or a page table for a graph whose columns stay in host
// memory (PagedCols: page e >> shift is a device cache slot or the pinned
// host array itself, each pointer rebased so that it takes the absolute edge
// position).  PagedCols also counts a 1-in-64 sample of the reads per page,
// the demand the host's page cache policy follows.
struct PlainCols {
    __device__ __forceinline__ uint32_t operator()(const Graph& g, uint64_t e, uint32_t, uint32_t) const {
        return g.col[e];
    }
};

struct PagedCols {
    const uint32_t* const* __restrict__ pages;
    uint32_t* __restrict__ hits;
    uint32_t shift;
    __device__ __forceinline__ uint32_t operator()(const Graph&, uint64_t e, uint32_t id, uint32_t hop) const {
        const uint64_t pg = e >> shift;
        if (((id + hop * 0x9E3779B9u) >> 26) == 0) atomicAdd(&hits[pg], 1u);
        return pages[pg][e];
    }
};

// One step of a walk (its state set up by start_walk).  Returns true when the
// walk has ended; the final vertex/hop stay in the state, as the reference
// leaves them in the record it drops.  Else-branches per kind follow advance()
// in gw_device.h.  The draw of the next hop depends only on (id, hop + 1), so
// it is computed while the edge load is in flight.
template <int KIND, bool EXT, class Rec, class Cols = PlainCols>
__device__ __forceinline__ bool step_once(const Prog& p, const Graph& g, State& s, Rec& rec, const Cols& cols = Cols{}) {
    constexpr bool capped = (KIND == GW_JUMP || KIND == GW_STOP || KIND == GW_RESTART);
    if (capped && s.hop >= p.L) return true;
    const u32x4 r = s.r;
    uint64_t b, deg;
    if (EXT) {
        b = s.row;
        deg = s.deg;
    } else {
        b = g.off[s.v];
        deg = g.off[s.v + 1] - b;
    }
    // The next edge's load is issued before the visit is recorded: a recorder
    // that waits on memory (a hash probe) then overlaps the edge's latency.
    // Counts commute and a walk's own visits keep their order.
    if (deg > 0 && static_cast<uint64_t>(r.x) >= p.stop_thr) {
        const uint64_t x = static_cast<uint64_t>(r.y) | (static_cast<uint64_t>(r.z) << 32);
        const uint64_t e = b + mulhi64(x, deg);
        if (EXT) {
            const Edge ed = g.ext[e];
            rec.visit(s.v, s.hop);
            s.r = philox4x32_10(s.id, s.hop + 1, kTagWalk, p.stream, p.k0, p.k1);
            s.v = ed.dst;
            s.deg = ed.deg;
            s.row = ed.row;
        } else {
            const uint32_t nv = cols(g, e, s.id, s.hop);
            rec.visit(s.v, s.hop);
            s.r = philox4x32_10(s.id, s.hop + 1, kTagWalk, p.stream, p.k0, p.k1);
            s.v = nv;
        }
    } else {
        rec.visit(s.v, s.hop);
        if (KIND == GW_STOP || KIND == GW_PROB) return true;
        s.r = philox4x32_10(s.id, s.hop + 1, kTagWalk, p.stream, p.k0, p.k1);
        if (KIND == GW_JUMP) {
            s.v = mulhi32(r.w, p.nvertices);
            load_row<EXT>(g, s);
        } else {
            s.v = s.source;
            s.row = s.srow;
            s.deg = s.sdeg;
        }
    }
    ++s.hop;
    ++rec.steps;
    if (KIND == GW_RESTART_JOINT) return s.hop % p.L == p.L - 1;
    if (capped) return s.hop >= p.L;
    return false;
}

// Wavefront-wide sum of a 32-bit value (64 lanes), result valid in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Adds a and b, summed over the workgroup, to d0 and d1: one atomic per
// workgroup and counter (none for a zero sum) instead of one per wavefront --
// a persistent grid's wavefronts all add to the same two addresses, where
// device-scope atomics serialise.  Every thread of the workgroup calls it.
__device__ __forceinline__ void workgroup_add2(unsigned long long* d0, unsigned long long* d1, uint32_t a, uint32_t b) {
    __shared__ unsigned long long s_tot[2];
    if (threadIdx.x == 0) {
        s_tot[0] = 0;
        s_tot[1] = 0;
    }
    __syncthreads();
    const uint32_t wa = wave_sum(a), wb = wave_sum(b);
    if (__lane_id() == 0) {
        if (wa) atomicAdd(&s_tot[0], static_cast<unsigned long long>(wa));
        if (wb) atomicAdd(&s_tot[1], static_cast<unsigned long long>(wb));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_tot[0]) atomicAdd(d0, s_tot[0]);
        if (s_tot[1]) atomicAdd(d1, s_tot[1]);
    }
}

// Block-wide top-k over n items with unique 64-bit rank keys (larger is
// better; key 0 marks an empty item and is never selected), k <= kTopkMax.
// An MSB radix select over LDS histograms finds the k-th largest key: one
// pass counts the items and finds the largest key, the byte passes start at
// its highest non-zero byte and stop as soon as the chosen bucket holds
// exactly the keys still needed.  The keys at or above the threshold are
// gathered into sel[] and (kSorted) sorted descending (by rank for at most
// 256 keys, else a bitonic network).  Returns
// min(k, non-empty items); every thread of the block calls it and gets the
// same value.
constexpr uint32_t kTopkMax = 1024;

struct TopkScratch {
    uint32_t hist[256];
    unsigned long long prefix, maxkey;
    uint32_t remaining, fill, valid;
    int shift_done;  // the select stopped after the pass at this shift (-1: all passes)
    unsigned long long sel[kTopkMax];
};

template <bool kSorted = true, uint32_t kCache = 0, class KeyAt>
__device__ uint32_t block_topk(uint64_t n, uint32_t k, KeyAt key_at, TopkScratch& t) {
    // With kCache, a block of at most kCache x blockDim items keeps its keys in
    // registers: one read per item instead of one per pass.  `each` visits a
    // thread's items in index order either way.
    constexpr uint32_t kC = kCache ? kCache : 1;
    unsigned long long kc[kC];
    const bool cached = kCache && n <= static_cast<uint64_t>(kCache) * blockDim.x;
    if (cached) {
#pragma unroll
        for (uint32_t j = 0; j < kC; ++j) {
            const uint64_t i = threadIdx.x + static_cast<uint64_t>(j) * blockDim.x;
            kc[j] = i < n ? key_at(i) : 0ull;
        }
    }
    auto each = [&](auto&& body) {
        if (cached) {
#pragma unroll
            for (uint32_t j = 0; j < kC; ++j)
                if (threadIdx.x + static_cast<uint64_t>(j) * blockDim.x < n) body(kc[j]);
        } else {
            for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) body(key_at(i));
        }
    };
    if (threadIdx.x == 0) {
        t.valid = 0;
        t.fill = 0;
        t.prefix = 0;
        t.maxkey = 0;
        t.remaining = k;
        t.shift_done = -1;
    }
    __syncthreads();
    uint32_t v = 0;
    unsigned long long mx = 0;
    each([&](unsigned long long key) {
        v += key != 0ull;
        mx = key > mx ? key : mx;
    });
    if (v) atomicAdd(&t.valid, v);
    if (mx) atomicMax(&t.maxkey, mx);
    __syncthreads();
    const uint32_t valid = t.valid;
    unsigned long long thr = 1;  // gather keys >= thr (every non-empty key when valid <= k)
    if (valid > k) {
        // bytes above the largest key's highest non-zero byte are zero in every key
        int top = 56;
        while (top > 0 && (t.maxkey >> top) == 0) top -= 8;
        for (int shift = top; shift >= 0; shift -= 8) {
            for (uint32_t b = threadIdx.x; b < 256; b += blockDim.x) t.hist[b] = 0;
            __syncthreads();
            const unsigned long long prefix = t.prefix;
            // Visit counts are skewed (most pairs hold 1 or 2), so most keys
            // share a bin: a thread folds runs of one bin in registers and adds
            // each run with one LDS atomic.
            uint32_t cur = 0, run = 0;
            each([&](unsigned long long key) {
                if (key != 0ull && (shift == top || (key >> (shift + 8)) == prefix)) {
                    const uint32_t d = static_cast<uint32_t>(key >> shift) & 255u;
                    if (d == cur) {
                        ++run;
                    } else {
                        if (run) atomicAdd(&t.hist[cur], run);
                        cur = d;
                        run = 1;
                    }
                }
            });
            if (run) atomicAdd(&t.hist[cur], run);
            __syncthreads();
            // The bin that holds the remaining-th largest key: the first wavefront
            // scans the 256 bins from the top, four per lane, and the one lane
            // whose bin crosses `remaining` records it.
            if (threadIdx.x < 64) {
                const uint32_t l = threadIdx.x;
                uint32_t h[4], s = 0;
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    h[j] = t.hist[255 - (4 * l + j)];
                    s += h[j];
                }
                uint32_t incl = s;
#pragma unroll
                for (uint32_t o = 1; o < 64; o <<= 1) {
                    const uint32_t y = __shfl_up(incl, o, 64);
                    if (l >= o) incl += y;
                }
                const uint32_t rem = t.remaining;
                uint32_t above = incl - s;
#pragma unroll
                for (uint32_t j = 0; j < 4; ++j) {
                    if (above < rem && above + h[j] >= rem) {
                        t.prefix = (prefix << 8) | static_cast<unsigned long long>(255 - (4 * l + j));
                        t.remaining = rem - above;
                        // the whole bucket is needed: its smallest possible key is the threshold
                        if (h[j] == rem - above) t.shift_done = shift;
                    }
                    above += h[j];
                }
            }
            __syncthreads();
            if (t.shift_done >= 0) break;
        }
        thr = t.shift_done >= 0 ? t.prefix << t.shift_done : t.prefix;
        if (thr == 0) thr = 1;
    }
    each([&](unsigned long long key) {
        if (key != 0ull && key >= thr) t.sel[atomicAdd(&t.fill, 1u)] = key;
    });
    __syncthreads();
    const uint32_t m = t.fill;
    if (!kSorted) return m;  // the caller takes the set, in no order
    if (m <= 256 && m <= blockDim.x) {
        // few keys: each goes straight to its rank (the keys above it; keys
        // are unique), read against every key while all threads read the
        // same one -- one barrier instead of a bitonic network's
        unsigned long long mine = 0;
        uint32_t rank = 0;
        if (threadIdx.x < m) {
            mine = t.sel[threadIdx.x];
            for (uint32_t j = 0; j < m; ++j) rank += t.sel[j] > mine;
        }
        __syncthreads();
        if (threadIdx.x < m) t.sel[rank] = mine;
        __syncthreads();
        return m;
    }
    uint32_t size = 1;
    while (size < m) size <<= 1;
    for (uint32_t i = m + threadIdx.x; i < size; i += blockDim.x) t.sel[i] = 0;
    __syncthreads();
    for (uint32_t len = 2; len <= size; len <<= 1) {
        for (uint32_t j = len >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) {
                const uint32_t q = i ^ j;
                if (q > i) {
                    const bool desc = (i & len) == 0;
                    const unsigned long long x = t.sel[i], y = t.sel[q];
                    if (desc ? x < y : x > y) {
                        t.sel[i] = y;
                        t.sel[q] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    return m;
}

// Rank key of a (vertex, count) pair: more visits first, then the smaller vertex.
__device__ __forceinline__ unsigned long long topk_key(uint32_t vertex, uint32_t count) {
    return count ? (static_cast<unsigned long long>(count) << 32) | static_cast<uint32_t>(~vertex) : 0ull;
}

}  // namespace gw
