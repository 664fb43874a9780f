// gw_table.hip -- aggregation and read-out of GW_REC_SOURCE_COUNT results.
//
//  * Global sort-reduce: the step kernel's (source, vertex) key log is
//    radix-sorted (rocPRIM through hipCUB) and run-length encoded.
//  * Partitioned: the key log is scattered into one region per source field
//    (vertex ids only), then one workgroup per source counts its region in an
//    LDS hash table.  A source whose distinct vertices outgrow the table
//    flushes partial pairs to a split list that is sort-reduced exactly.
//  * Hash: non-empty slots of the device hash table are compacted and sorted
//    at read-out.
// Read-out unpacks (source, vertex, count) triples sorted by source, vertex.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "gw_device.h"
#include "gw_internal.h"

namespace {

// Occupied slots of the hash table -> (key, count) arrays, one atomic per
// wavefront (ballot ranking).  Packed slots hold (key << cbits) | count; wide
// slots hold {key + 1, count}.
__global__ void compact_kernel(const unsigned long long* __restrict__ tab, uint64_t slots, uint32_t wide,
                               uint32_t cbits, unsigned long long* __restrict__ out_k, uint32_t* __restrict__ out_c,
                               unsigned long long* __restrict__ n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    // workgroup-uniform trip count: every lane of a wavefront takes part in the ballots
    for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); i0 < slots; i0 += stride) {
        const uint64_t i = i0 + threadIdx.x;
        unsigned long long k = 0, c = 0;
        if (i < slots) {
            if (wide) {
                k = tab[2 * i];
                c = tab[2 * i + 1];
            } else {
                const unsigned long long x = tab[i];
                k = x ? (x >> cbits) + 1 : 0ull;  // + 1: 0 stays "empty"
                c = x & ((1ull << cbits) - 1);
            }
        }
        const bool full = k != 0ull;
        const uint64_t mask = __ballot(full);
        if (mask == 0) continue;
        const uint32_t lane = __lane_id();
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(n, static_cast<unsigned long long>(__popcll(mask)));
        base = gw::lane_value(base, leader);
        if (full) {
            const uint32_t r = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
            out_k[base + r] = k - 1;
            out_c[base + r] = static_cast<uint32_t>(c);
        }
    }
}

__global__ void unpack_runs_kernel(const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ counts,
                                   uint64_t n, uint32_t vbits, uint32_t* __restrict__ s, uint32_t* __restrict__ v,
                                   uint32_t* __restrict__ c) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
        const unsigned long long k = keys[i];
        v[i] = static_cast<uint32_t>(k & ((1ull << vbits) - 1));
        s[i] = static_cast<uint32_t>(k >> vbits);
        c[i] = counts[i];
    }
}


constexpr uint32_t kAggLog2 = 14;
constexpr uint32_t kAggSlots = 1u << kAggLog2;               // LDS hash slots per workgroup (key + count: 128 KiB)
constexpr uint32_t kAggThreads = 1024;
constexpr uint32_t kAggPerThread = 4;
constexpr uint32_t kAggTile = kAggThreads * kAggPerThread;   // keys loaded per iteration
constexpr uint32_t kAggWaves = kAggThreads / 64;
constexpr uint32_t kAggHeld = 12;                             // keys a thread holds for a whole source
constexpr uint32_t kAggHeldKeys = kAggHeld * kAggThreads;
static_assert(kAggHeldKeys <= kAggSlots / 4 * 3, "held keys must fit a table that never flushes");
// A flushed table takes a whole tile plus the designated vertex below its 3/4 fill bound.
static_assert(kAggTile + 1 <= kAggSlots / 4 * 3, "a tile must fit an emptied table");
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;                   // never a vertex id (n <= 2^32 - 1)
constexpr uint32_t kPartPerLane = 4;                         // log slots per lane per partition sweep

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Zero the engine's device counters and the per-source region cursors before
// a run, in one launch.
__global__ void reset_counters_kernel(unsigned long long* __restrict__ counters, uint32_t ncounters,
                                      unsigned long long* __restrict__ region_n, uint32_t* __restrict__ fcount,
                                      uint32_t nf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncounters) counters[i] = 0ull;
    if (i < nf) {
        region_n[i] = 0ull;
        fcount[i] = 0u;
    }
}

// Scatter the key log into per-source regions: region f receives the vertex
// ids of source f's visits in [region_off[f], region_off[f + 1]).  Each
// wavefront sweeps 256 consecutive log slots at a time (4 per lane); the
// slots are grouped by source with a ballot loop, one atomic per group.  The
// slots of a wavefront-sized stretch of the log come from one step-kernel
// wavefront and name one to three sources, so a sweep costs one to three
// atomics.  The log length is read from the device (*n_ptr, capped at `cap`),
// so no host round trip separates the step kernel from this one.  A key that
// does not fit its region, or names a field >= nf, counts in `overflow` and
// the host discards the partition.
__global__ __launch_bounds__(256) void partition_kernel(const unsigned long long* __restrict__ log,
                                                       const unsigned long long* __restrict__ n_ptr, uint64_t cap,
                                                       uint64_t sentinel, uint32_t vbits, uint32_t nf,
                                                       const uint64_t* __restrict__ region_off,
                                                       unsigned long long* __restrict__ region_n,
                                                       uint32_t* __restrict__ out,
                                                       unsigned long long* __restrict__ overflow) {
    const uint64_t n = std::min<uint64_t>(*n_ptr, cap);
    const uint32_t lane = __lane_id();
    const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
    const uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
    constexpr uint32_t kSweep = 64 * kPartPerLane;
    for (uint64_t base = wave * kSweep; base < n; base += nwaves * kSweep) {
        uint32_t f[kPartPerLane], v[kPartPerLane];
        bool pend[kPartPerLane];
#pragma unroll
        for (uint32_t j = 0; j < kPartPerLane; ++j) {
            const uint64_t i = base + j * 64 + lane;
            const uint64_t key = i < n ? log[i] : sentinel;
            pend[j] = key != sentinel;
            f[j] = static_cast<uint32_t>(key >> vbits);
            v[j] = static_cast<uint32_t>(key & ((1ull << vbits) - 1));
            if (pend[j] && f[j] >= nf) {
                atomicAdd(overflow, 1ull);
                pend[j] = false;
            }
        }
        for (;;) {
            // leader: the first pending slot of the sweep (wave-uniform search)
            bool any = false;
            uint32_t fl = 0;
#pragma unroll
            for (uint32_t j = 0; j < kPartPerLane; ++j) {
                const uint64_t b = __ballot(pend[j]);
                if (!any && b) {
                    fl = __shfl(f[j], static_cast<int>(__ffsll(static_cast<unsigned long long>(b)) - 1), 64);
                    any = true;
                }
            }
            if (!any) break;
            uint64_t mt[kPartPerLane];
            uint32_t tot = 0;
#pragma unroll
            for (uint32_t j = 0; j < kPartPerLane; ++j) {
                mt[j] = __ballot(pend[j] && f[j] == fl);
                tot += __popcll(mt[j]);
            }
            unsigned long long at = 0;
            if (lane == 0) at = atomicAdd(&region_n[fl], static_cast<unsigned long long>(tot));
            at = gw::lane_value(at, 0u);
            const uint64_t lo = region_off[fl], hi = region_off[fl + 1];
#pragma unroll
            for (uint32_t j = 0; j < kPartPerLane; ++j) {
                if (pend[j] && f[j] == fl) {
                    const uint64_t pos = lo + at + rank_in(mt[j]);
                    if (pos < hi) out[pos] = v[j];
                    else atomicAdd(overflow, 1ull);
                    pend[j] = false;
                }
                at += __popcll(mt[j]);
            }
        }
    }
}

// Table size for a region of n keys: the smallest power of two with n at
// most 3/4 of it (so a region that fits never flushes), within
// [2^10, kAggSlots].
__device__ __forceinline__ uint32_t agg_log2(uint64_t n) {
    const uint64_t want = n + n / 3 + 1;
    uint32_t lg = 10;
    while (lg < kAggLog2 && (1ull << lg) < want) ++lg;
    return lg;
}

// Write the occupied slots of the table (first `tslots`) from position `at0`
// on, then clear them.  Wavefront w owns the slots [w * tslots / 16, (w + 1)
// * tslots / 16): it counts its occupied slots, the counts meet in s_wtot,
// and each wavefront writes its pairs in slot order after those of the
// wavefronts before it (no atomics).  kSplit: as 64-bit keys (fbase |
// vertex) + counts into the split list (only if `fits`); else as vertex +
// count into the source's own region.  Starts after the inserts have been
// synchronised; ends synchronised.
template <bool kSplit>
__device__ __forceinline__ void agg_write(uint32_t* s_key, uint32_t* s_cnt, uint32_t tslots, uint32_t* s_wtot,
                                          uint64_t at0, bool fits, uint64_t fbase, uint32_t* __restrict__ seg_v,
                                          uint32_t* __restrict__ seg_c, unsigned long long* __restrict__ split_k,
                                          uint32_t* __restrict__ split_c) {
    const uint32_t lane = __lane_id(), w = threadIdx.x / 64;
    const uint32_t spw = tslots / kAggWaves;  // tslots >= 1024: a multiple of 64, every lane present
    const uint32_t s0 = w * spw;
    uint32_t occ = 0;
    for (uint32_t s = lane; s < spw; s += 64) occ += __popcll(__ballot(s_key[s0 + s] != kEmptyKey));
    if (lane == 0) s_wtot[w] = occ;
    __syncthreads();
    uint64_t at = at0 + gw::wave_sum(lane < w ? s_wtot[lane] : 0u);
    for (uint32_t s = lane; s < spw; s += 64) {
        const uint32_t i = s0 + s;
        const uint32_t k = s_key[i];
        const bool full = k != kEmptyKey;
        const uint64_t mask = __ballot(full);
        if (full && fits) {
            const uint64_t pos = at + rank_in(mask);
            const uint32_t c = s_cnt[i];
            if (kSplit) {
                split_k[pos] = fbase | k;
                split_c[pos] = c;
            } else {
                seg_v[pos] = k;
                seg_c[pos] = c;
            }
        }
        at += __popcll(mask);
        s_key[i] = kEmptyKey;
        s_cnt[i] = 0;
    }
    __syncthreads();
}

// Flush the table (`total` occupied slots) to the split list: one returning
// atomic reserves the pairs (counted even when they do not fit `cap`; the
// host checks).
__device__ __forceinline__ void agg_flush(uint32_t* s_key, uint32_t* s_cnt, uint32_t tslots, uint32_t total,
                                          uint32_t* s_wtot, unsigned long long* s_base, uint64_t fbase,
                                          unsigned long long* __restrict__ split_k, uint32_t* __restrict__ split_c,
                                          unsigned long long* __restrict__ split_n, uint64_t cap) {
    if (threadIdx.x == 0) *s_base = total ? atomicAdd(split_n, static_cast<unsigned long long>(total)) : 0ull;
    __syncthreads();
    const uint64_t b = *s_base;
    agg_write<true>(s_key, s_cnt, tslots, s_wtot, b, b + total <= cap, fbase, nullptr, nullptr, split_k, split_c);
}

// Source f's region [lo, lo + n) and, in direct-region runs (fcount set), its
// designated vertex's count fc and id fv (fv = kEmptyKey when fc is 0).
struct AggMeta {
    uint64_t lo = 0, n = 0;
    uint32_t fc = 0, fv = kEmptyKey;
};

__device__ __forceinline__ AggMeta region_meta(const uint64_t* __restrict__ region_off,
                                               const unsigned long long* __restrict__ region_n, uint32_t nf,
                                               uint32_t f, const uint32_t* __restrict__ fvert,
                                               const uint32_t* __restrict__ fcount) {
    AggMeta m;
    if (f < nf) {
        m.lo = region_off[f];
        // clamp to the region: an overflowed partition is discarded by the
        // host, but must still be read in bounds
        m.n = std::min<uint64_t>(region_n[f], region_off[f + 1] - m.lo);
        if (fcount) {
            const uint32_t c = fcount[f], v = fvert[f];
            m.fc = c;
            m.fv = c ? v : kEmptyKey;
        }
    }
    return m;
}

// A source with region keys, as the aggregation workgroups take it.
struct alignas(16) AggItem {
    uint64_t lo, n;
    uint32_t f, fc, fv, pad;
};

// One thread per source.  A source with no region keys is finished here --
// its designated vertex's pair (if it has visits) written over the start of
// its region, pairs[f] = 0 or 1 -- and every other source is appended to the
// work list of part_aggregate_kernel (one atomic per wavefront).  Most
// sources of a batch can be of the first kind (walks from a vertex with no
// out-edges end at once), and they then cost the workgroups nothing.
__global__ __launch_bounds__(256) void part_classify_kernel(
    uint32_t* __restrict__ keys, uint32_t* __restrict__ counts, unsigned long long* __restrict__ pairs,
    const uint64_t* __restrict__ region_off, const unsigned long long* __restrict__ region_n, uint32_t nf,
    unsigned long long* __restrict__ total_pairs, AggItem* __restrict__ items, unsigned long long* __restrict__ nitems,
    const uint32_t* __restrict__ fvert, const uint32_t* __restrict__ fcount) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = __lane_id();
    const AggMeta m = region_meta(region_off, region_n, nf, f, fvert, fcount);
    const bool work = f < nf && m.n > 0;
    uint32_t single = 0;
    if (f < nf && !work) {
        if (m.fc) {
            keys[m.lo] = m.fv;
            counts[m.lo] = m.fc;
            single = 1;
        }
        pairs[f] = single;
    }
    const uint64_t mask = __ballot(work);
    if (mask) {
        unsigned long long at = 0;
        if (lane == 0) at = atomicAdd(nitems, static_cast<unsigned long long>(__popcll(mask)));
        at = gw::lane_value(at, 0u);
        if (work) items[at + rank_in(mask)] = AggItem{m.lo, m.n, f, m.fc, m.fv, 0u};
    }
    single = gw::wave_sum(single);
    if (lane == 0 && single) atomicAdd(total_pairs, static_cast<unsigned long long>(single));
}

__device__ __forceinline__ AggItem agg_item(const AggItem* __restrict__ items, uint64_t nitems, uint64_t i) {
    return i < nitems ? items[i] : AggItem{0, 0, 0, 0, kEmptyKey, 0};
}

// A source whose keys all fit in registers and in a table it never has to
// flush (kAggHeld keys a thread; with its designated vertex at most 3/4 of
// kAggSlots).
__device__ __forceinline__ bool agg_held(const AggItem& m) {
    return m.n + (m.fc ? 1u : 0u) <= kAggHeldKeys;
}

__device__ __forceinline__ void load_held(const uint32_t* __restrict__ keys, const AggItem& m,
                                          uint32_t (&t)[kAggHeld]) {
    const bool held = agg_held(m);
#pragma unroll
    for (uint32_t j = 0; j < kAggHeld; ++j) {
        const uint64_t idx = j * kAggThreads + threadIdx.x;
        t[j] = held && idx < m.n ? keys[m.lo + idx] : kEmptyKey;
    }
}

__device__ __forceinline__ void load_tile(const uint32_t* __restrict__ keys, uint64_t lo, uint64_t n, uint64_t first,
                                          uint32_t (&t)[kAggPerThread]) {
#pragma unroll
    for (uint32_t j = 0; j < kAggPerThread; ++j) {
        const uint64_t idx = first + j * kAggThreads + threadIdx.x;
        t[j] = idx < n ? keys[lo + idx] : kEmptyKey;
    }
}

__device__ __forceinline__ uint32_t agg_home(uint32_t v, uint32_t lg) { return (v * 0x9E3779B1u) >> (32 - lg); }

// Adds `inc` visits of v to the table (linear probing from slot h, v's home
// slot or a later one of its probe sequence); 1 if v took an empty slot.
__device__ __forceinline__ uint32_t agg_put(uint32_t* s_key, uint32_t* s_cnt, uint32_t lg, uint32_t v, uint32_t inc,
                                            uint32_t h) {
    for (;;) {
        const uint32_t old = atomicCAS(&s_key[h], kEmptyKey, v);
        if (old == kEmptyKey || old == v) {
            atomicAdd(&s_cnt[h], inc);
            return old == kEmptyKey;
        }
        h = (h + 1) & ((1u << lg) - 1);
    }
}

// One key a lane (kEmptyKey = none).  Runs of one vertex across the wavefront
// (the first hops of a source's walks) are combined before touching LDS: the
// lowest lane holding the first valid lane's vertex adds the whole group.
// (Measured alternatives that were slower: probing the lane's keys
// together, 4-slot buckets read with one 128-bit load, and 64-bit key|count
// slots read before the atomic; plain stores of every key to its home slot
// read back after a barrier, with compare-and-swap only for the keys that
// lost their slot; a lane's keys probed in batches of compare-and-swaps
// issued back to back.)
__device__ __forceinline__ uint32_t agg_insert(uint32_t* s_key, uint32_t* s_cnt, uint32_t lg, uint32_t v,
                                               uint32_t h) {
    const uint64_t valid = __ballot(v != kEmptyKey);
    if (!valid) return 0;  // wavefront-uniform
    uint32_t inc = 1;
    const uint32_t v0 = __shfl(v, static_cast<int>(__ffsll(static_cast<unsigned long long>(valid)) - 1), 64);
    const uint64_t same = __ballot(v == v0);
    if (v == v0) {
        if (rank_in(same) == 0) inc = __popcll(same);
        else v = kEmptyKey;
    }
    return v == kEmptyKey ? 0u : agg_put(s_key, s_cnt, lg, v, inc, h);
}

// Persistent workgroups, one per CU, each counting one source of the work
// list at a time in an LDS open-addressing table sized to the source's key
// count; workgroup b takes the items b, b + G, b + 2G, ... (G workgroups:
// the sources of a batch are alike in size, and a fixed order needs no
// queue atomics, whose returns would stall the pipeline below).  Right
// after a source's keys are in the table, the workgroup issues the loads of
// the next source's keys (into registers) and of the list entry after that,
// so they are in flight while the table is written out, and a source's
// inserts wait only on loads issued a phase earlier.  A source with more
// keys than a workgroup holds (kAggHeldKeys) goes tile by tile: before a
// tile would lift the fill above 3/4 of the table, the table is flushed to
// the split list, so every probe sequence ends.  A source that never flushed
// writes its distinct vertices and counts over the start of its own region
// (its keys are all in registers or consumed by then) and its pair count to
// pairs[f]; a flushed source emits partial pairs to the split list, summed
// later by sort-reduce, and pairs[f] = 0.  A source's designated vertex
// (direct-region runs) and its count fc, counted apart by the step kernel,
// go into the table with the region's keys.
__global__ __launch_bounds__(kAggThreads) void part_aggregate_kernel(
    const AggItem* __restrict__ items, const unsigned long long* __restrict__ nitems_p, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ counts, unsigned long long* __restrict__ pairs, uint32_t vbits,
    unsigned long long* __restrict__ total_pairs, unsigned long long* __restrict__ split_k,
    uint32_t* __restrict__ split_c, unsigned long long* __restrict__ split_n, uint64_t split_cap,
    unsigned long long* __restrict__ flushed_sources) {
    __shared__ uint32_t s_key[kAggSlots];
    __shared__ uint32_t s_cnt[kAggSlots];
    __shared__ uint32_t s_wtot[kAggWaves];
    __shared__ uint32_t s_fill;
    __shared__ unsigned long long s_base;
    const uint64_t nitems = *nitems_p;
    if (blockIdx.x >= nitems) return;  // workgroup-uniform, before any barrier
    const uint64_t G = gridDim.x;
    const uint32_t lane = __lane_id();
    if (threadIdx.x == 0) s_fill = 0;
    for (uint32_t i = threadIdx.x; i < kAggSlots; i += kAggThreads) {
        s_key[i] = kEmptyKey;
        s_cnt[i] = 0;
    }
    __syncthreads();
    uint64_t x = blockIdx.x;
    AggItem cm = agg_item(items, nitems, x);
    AggItem nm = agg_item(items, nitems, x + G);
    uint32_t cur[kAggHeld], nxt[kAggHeld];
    load_held(keys, cm, cur);
    while (x < nitems) {
        const uint64_t lo = cm.lo, n = cm.n;
        const uint32_t f = cm.f, fc = cm.fc, fv = cm.fv;
        const uint32_t lg = agg_log2(n + (fc ? 1u : 0u));
        const uint32_t tslots = 1u << lg;
        AggItem nnm;
        if (agg_held(cm)) {
            uint32_t added = 0;
            if (threadIdx.x == 0 && fc) added = agg_put(s_key, s_cnt, lg, fv, fc, agg_home(fv, lg));
#pragma unroll
            for (uint32_t j = 0; j < kAggHeld; ++j)
                if (j * kAggThreads < n) added += agg_insert(s_key, s_cnt, lg, cur[j], agg_home(cur[j], lg));  // uniform
            added = gw::wave_sum(added);
            if (lane == 0 && added) atomicAdd(&s_fill, added);
            __syncthreads();
            const uint32_t total = s_fill;
            nnm = agg_item(items, nitems, x + 2 * G);
            load_held(keys, nm, nxt);
            agg_write<false>(s_key, s_cnt, tslots, s_wtot, lo, true, 0, keys, counts, nullptr, nullptr);
            if (threadIdx.x == 0) {
                pairs[f] = total;
                atomicAdd(total_pairs, static_cast<unsigned long long>(total));
                s_fill = 0;
            }
        } else {
            const uint32_t max_fill = tslots / 4 * 3;
            const uint64_t fbase = static_cast<uint64_t>(f) << vbits;
            bool flushed = false;
            if (fc) {  // workgroup-uniform: the table is empty when a source starts
                if (threadIdx.x == 0) {
                    agg_put(s_key, s_cnt, lg, fv, fc, agg_home(fv, lg));
                    s_fill = 1;
                }
                __syncthreads();
            }
            uint32_t tc[kAggPerThread], tp[kAggPerThread];
            load_tile(keys, lo, n, 0, tc);
            uint32_t fill = 0;
            for (uint64_t base = 0;; base += kAggTile) {
                const uint64_t next = base + kAggTile;
                if (next < n) load_tile(keys, lo, n, next, tp);
                uint32_t added = 0;
#pragma unroll
                for (uint32_t j = 0; j < kAggPerThread; ++j) added += agg_insert(s_key, s_cnt, lg, tc[j], agg_home(tc[j], lg));
                added = gw::wave_sum(added);
                if (lane == 0 && added) atomicAdd(&s_fill, added);
                __syncthreads();
                fill = s_fill;
                __syncthreads();  // every thread has read the fill before it is reset
                if (next >= n) break;
                if (fill + std::min<uint64_t>(kAggTile, n - next) > max_fill) {  // workgroup-uniform
                    agg_flush(s_key, s_cnt, tslots, fill, s_wtot, &s_base, fbase, split_k, split_c, split_n,
                              split_cap);
                    if (threadIdx.x == 0) s_fill = 0;
                    flushed = true;
                    __syncthreads();
                }
#pragma unroll
                for (uint32_t j = 0; j < kAggPerThread; ++j) tc[j] = tp[j];
            }
            if (flushed) {
                agg_flush(s_key, s_cnt, tslots, fill, s_wtot, &s_base, fbase, split_k, split_c, split_n, split_cap);
                if (threadIdx.x == 0) {
                    pairs[f] = 0;
                    atomicAdd(flushed_sources, 1ull);
                }
            } else {
                agg_write<false>(s_key, s_cnt, tslots, s_wtot, lo, true, 0, keys, counts, nullptr, nullptr);
                if (threadIdx.x == 0) {
                    pairs[f] = fill;
                    atomicAdd(total_pairs, static_cast<unsigned long long>(fill));
                }
            }
            if (threadIdx.x == 0) s_fill = 0;
            nnm = agg_item(items, nitems, x + 2 * G);
            load_held(keys, nm, nxt);
        }
        x += G;
        cm = nm;
        nm = nnm;
#pragma unroll
        for (uint32_t j = 0; j < kAggHeld; ++j) cur[j] = nxt[j];
    }
}

// Read-out of the partitioned result: the pairs at the start of each source's
// region are packed into (fbase | vertex, count) at a cursor, one workgroup
// per source (order is restored by the sort that follows).
__global__ __launch_bounds__(256) void gather_regions_kernel(const uint32_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ counts,
                                                            const unsigned long long* __restrict__ pairs,
                                                            const uint64_t* __restrict__ region_off, uint32_t vbits,
                                                            unsigned long long* __restrict__ out_k,
                                                            uint32_t* __restrict__ out_c,
                                                            unsigned long long* __restrict__ cursor) {
    __shared__ unsigned long long s_at;
    const uint32_t f = blockIdx.x;
    const uint64_t np = pairs[f];
    if (np == 0) return;
    if (threadIdx.x == 0) s_at = atomicAdd(cursor, static_cast<unsigned long long>(np));
    __syncthreads();
    const uint64_t lo = region_off[f], at = s_at;
    const uint64_t fbase = static_cast<uint64_t>(f) << vbits;
    for (uint64_t i = threadIdx.x; i < np; i += blockDim.x) {
        out_k[at + i] = fbase | keys[lo + i];
        out_c[at + i] = counts[lo + i];
    }
}


constexpr uint32_t kTopkThreads = 1024;
using gw::kTopkMax;

__global__ void source_of_kernel(const unsigned long long* __restrict__ keys, uint64_t n, uint32_t vbits,
                                 uint64_t* __restrict__ src) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
        src[i] = keys[i] >> vbits;
}

// One workgroup per source segment [seg[g], seg[g+1]) of the sorted pairs:
// block_topk over the rank keys of the segment's (vertex, count) pairs.
__global__ __launch_bounds__(kTopkThreads) void topk_kernel(const unsigned long long* __restrict__ keys,
                                                           const uint32_t* __restrict__ counts,
                                                           const uint64_t* __restrict__ seg, uint32_t vbits,
                                                           uint32_t k, uint32_t* __restrict__ out_s,
                                                           uint32_t* __restrict__ out_v, uint32_t* __restrict__ out_c,
                                                           uint32_t* __restrict__ out_n) {
    __shared__ gw::TopkScratch t;
    const uint32_t g = blockIdx.x;
    const uint64_t lo = seg[g], n = seg[g + 1] - lo;
    const unsigned long long vmask = (1ull << vbits) - 1;
    const uint32_t m = gw::block_topk(
        n, k, [&](uint64_t i) { return gw::topk_key(static_cast<uint32_t>(keys[lo + i] & vmask), counts[lo + i]); }, t);
    const uint32_t src = static_cast<uint32_t>(keys[lo] >> vbits);
    for (uint32_t r = threadIdx.x; r < m; r += blockDim.x) {
        const unsigned long long key = t.sel[r];
        out_s[static_cast<uint64_t>(g) * k + r] = src;
        out_v[static_cast<uint64_t>(g) * k + r] = ~static_cast<uint32_t>(key);
        out_c[static_cast<uint64_t>(g) * k + r] = static_cast<uint32_t>(key >> 32);
    }
    if (threadIdx.x == 0) out_n[g] = m;
}

}  // namespace

// Sort-reduce of the key log: radix sort of the n logged slots on the key
// bits plus one (the sentinel bit), then run-length encoding; the runs are
// the (source, vertex) pairs in order with their visit counts.  The trailing
// sentinel run (slab padding), if any, is dropped.
static int ensure_runs(gw_ctx* ctx, uint64_t n) {
    if (n <= ctx->runs_cap) return GW_OK;
    gw_free(ctx, ctx->d_keys_alt, ctx->runs_cap * 8);
    gw_free(ctx, ctx->d_runs, ctx->runs_cap * 8);
    gw_free(ctx, ctx->d_run_counts, ctx->runs_cap * 4);
    ctx->d_keys_alt = ctx->d_runs = nullptr;
    ctx->d_run_counts = nullptr;
    ctx->runs_cap = 0;
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_keys_alt, n * 8, "key sort")) ||
        (rc = gw_alloc(ctx, (void**)&ctx->d_runs, n * 8, "key runs")) ||
        (rc = gw_alloc(ctx, (void**)&ctx->d_run_counts, n * 4, "key run counts")))
        return rc;
    ctx->runs_cap = n;
    return GW_OK;
}

static int ensure_sort_temp(gw_ctx* ctx, size_t need) {
    if (need <= ctx->sort_temp_bytes) return GW_OK;
    gw_free(ctx, ctx->d_sort_temp, ctx->sort_temp_bytes);
    ctx->d_sort_temp = nullptr;
    ctx->sort_temp_bytes = 0;
    int rc = gw_alloc(ctx, &ctx->d_sort_temp, need, "key sort temp");
    if (rc) return rc;
    ctx->sort_temp_bytes = need;
    return GW_OK;
}

int gw_keylog_reduce(gw_ctx* ctx, uint64_t n, double* ms) {
    const int end_bit = static_cast<int>(64 - ctx->table_cbits) + 1;
    int rc0 = ensure_runs(ctx, n);
    if (rc0) return rc0;
    size_t need_sort = 0, need_rle = 0;
    GW_HIP(ctx, rocprim::radix_sort_keys(nullptr, need_sort, ctx->d_keys, ctx->d_keys_alt, n, 0, end_bit,
                                                  ctx->stream));
    GW_HIP(ctx, rocprim::run_length_encode(nullptr, need_rle, ctx->d_keys_alt, n, ctx->d_runs, ctx->d_run_counts, ctx->d_counters + kCntRuns, ctx->stream));
    if ((rc0 = ensure_sort_temp(ctx, std::max(need_sort, need_rle)))) return rc0;
    size_t tb = ctx->sort_temp_bytes;
    GW_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    if (n) {
        GW_HIP(ctx, rocprim::radix_sort_keys(ctx->d_sort_temp, tb, ctx->d_keys, ctx->d_keys_alt, n, 0, end_bit,
                                                      ctx->stream));
        tb = ctx->sort_temp_bytes;
        GW_HIP(ctx, rocprim::run_length_encode(ctx->d_sort_temp, tb, ctx->d_keys_alt, n, ctx->d_runs, ctx->d_run_counts, ctx->d_counters + kCntRuns, ctx->stream));
    } else {
        GW_HIP(ctx, hipMemsetAsync(ctx->d_counters + kCntRuns, 0, 8, ctx->stream));
    }
    GW_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    unsigned long long runs = 0;
    GW_HIP(ctx, hipMemcpyAsync(&runs, ctx->d_counters + kCntRuns, 8, hipMemcpyDeviceToHost, ctx->stream));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float t = 0.f;
    GW_HIP(ctx, hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]));
    if (runs) {
        unsigned long long last = 0;
        GW_HIP(ctx, hipMemcpy(&last, ctx->d_runs + runs - 1, 8, hipMemcpyDeviceToHost));
        if (last == ctx->keys_sentinel) --runs;
    }
    ctx->n_runs = runs;
    if (ms) *ms = t;
    return GW_OK;
}

int gw_partlog_prepare(gw_ctx* ctx, uint32_t per_walk, bool* ok) {
    *ok = false;
    const uint32_t nf = static_cast<uint32_t>(ctx->field_walks.size());
    if (nf == 0) return GW_OK;
    int rc;
    // Regions sized for per_walk visits of every walk of a field.  Counts are
    // 32-bit, so a region stays below 2^32 keys.
    if (ctx->part_per_walk != per_walk || ctx->part_fields != nf) {
        std::vector<uint64_t> off(nf + 1ull, 0);
        for (uint32_t f = 0; f < nf; ++f) {
            if (ctx->field_walks[f] * per_walk >= (1ull << 32)) return GW_OK;
            off[f + 1] = off[f] + ctx->field_walks[f] * per_walk;
        }
        if (nf > ctx->part_fields_cap) {
            gw_free(ctx, ctx->d_part_off, (ctx->part_fields_cap + 1ull) * 8);
            gw_free(ctx, ctx->d_part_cursor, ctx->part_fields_cap * 8ull);
            gw_free(ctx, ctx->d_part_pairs, ctx->part_fields_cap * 8ull);
            gw_free(ctx, ctx->d_part_fvert, ctx->part_fields_cap * 4ull);
            gw_free(ctx, ctx->d_part_fcount, ctx->part_fields_cap * 4ull);
            gw_free(ctx, ctx->d_part_items, ctx->part_fields_cap * 32ull);
            ctx->d_part_off = nullptr;
            ctx->d_part_cursor = ctx->d_part_pairs = nullptr;
            ctx->d_part_fvert = ctx->d_part_fcount = nullptr;
            ctx->d_part_items = nullptr;
            ctx->part_fields_cap = 0;
            if ((rc = gw_alloc(ctx, (void**)&ctx->d_part_off, (nf + 1ull) * 8, "source regions")) ||
                (rc = gw_alloc(ctx, (void**)&ctx->d_part_cursor, nf * 8ull, "source region cursors")) ||
                (rc = gw_alloc(ctx, (void**)&ctx->d_part_pairs, nf * 8ull, "source pair counts")) ||
                (rc = gw_alloc(ctx, (void**)&ctx->d_part_fvert, nf * 4ull, "source field vertices")) ||
                (rc = gw_alloc(ctx, (void**)&ctx->d_part_fcount, nf * 4ull, "source field vertex counts")) ||
                (rc = gw_alloc(ctx, &ctx->d_part_items, nf * 32ull, "aggregation work list")))
                return rc;
            ctx->part_fields_cap = nf;
        }
        const uint64_t need = std::max<uint64_t>(1, off[nf]);
        if (need > ctx->part_keys_cap) {
            gw_free(ctx, ctx->d_part_keys, ctx->part_keys_cap * 4);
            gw_free(ctx, ctx->d_part_counts, ctx->part_keys_cap * 4);
            ctx->d_part_keys = ctx->d_part_counts = nullptr;
            ctx->part_keys_cap = 0;
            if (gw_alloc(ctx, (void**)&ctx->d_part_keys, need * 4, "partitioned key log") ||
                gw_alloc(ctx, (void**)&ctx->d_part_counts, need * 4, "partitioned counts")) {
                gw_free(ctx, ctx->d_part_keys, need * 4);
                ctx->d_part_keys = nullptr;
                ctx->err.clear();
                return GW_OK;  // does not fit: the global sort-reduce runs instead
            }
            ctx->part_keys_cap = need;
        }
        GW_HIP(ctx, hipMemcpy(ctx->d_part_off, off.data(), (nf + 1ull) * 8, hipMemcpyHostToDevice));
        if (ctx->field_vertex.size() != nf) return gw_fail(ctx, GW_ESTATE, "gw_run: source field table out of date");
        GW_HIP(ctx, hipMemcpy(ctx->d_part_fvert, ctx->field_vertex.data(), nf * 4ull, hipMemcpyHostToDevice));
        ctx->part_per_walk = per_walk;
        ctx->part_fields = nf;
        ctx->part_total = off[nf];
    }
    // Split pairs (flushed sources) go to the key log's memory -- consumed by
    // the partition by then, or never written when the step kernel fills the
    // regions directly -- with their counts here: at most one pair per visit.
    if (ctx->keys_cap > ctx->split_cap) {
        gw_free(ctx, ctx->d_split_counts, ctx->split_cap * 8);
        ctx->d_split_counts = nullptr;
        ctx->split_cap = 0;
        if ((rc = gw_alloc(ctx, (void**)&ctx->d_split_counts, ctx->keys_cap * 8, "split counts"))) return rc;
        ctx->split_cap = ctx->keys_cap;
    }
    *ok = true;
    return GW_OK;
}

int gw_partlog_reset(gw_ctx* ctx, bool part) {
    const uint32_t nf = part ? ctx->part_fields : 0u;
    const uint32_t n = std::max<uint32_t>(kCounters, nf);
    reset_counters_kernel<<<dim3((n + 255) / 256), dim3(256), 0, ctx->stream>>>(ctx->d_counters, kCounters,
                                                                                ctx->d_part_cursor,
                                                                                ctx->d_part_fcount, nf);
    GW_HIP(ctx, hipGetLastError());
    return GW_OK;
}

int gw_partlog_enqueue(gw_ctx* ctx, bool partition) {
    unsigned long long* c = ctx->d_counters;
    const uint32_t cus = std::max(1u, ctx->cu_count);
    if (partition) {  // else the step kernel filled the regions itself
        partition_kernel<<<dim3(cus * 8u), dim3(256), 0, ctx->stream>>>(
            ctx->d_keys, c + kCntLogCursor, ctx->keys_cap, ctx->keys_sentinel, ctx->table_vbits, ctx->part_fields,
            ctx->d_part_off, ctx->d_part_cursor, ctx->d_part_keys, c + kCntPartOverflow);
        GW_HIP(ctx, hipGetLastError());
    }
    const uint32_t nf = ctx->part_fields;
    AggItem* items = static_cast<AggItem*>(ctx->d_part_items);
    part_classify_kernel<<<dim3((nf + 255) / 256), dim3(256), 0, ctx->stream>>>(
        ctx->d_part_keys, ctx->d_part_counts, ctx->d_part_pairs, ctx->d_part_off, ctx->d_part_cursor, nf,
        c + kCntOutPairs, items, c + kCntAggItems, partition ? nullptr : ctx->d_part_fvert,
        partition ? nullptr : ctx->d_part_fcount);
    GW_HIP(ctx, hipGetLastError());
    // workgroups beyond the work list's length return at once
    part_aggregate_kernel<<<dim3(std::min(nf, cus)), dim3(kAggThreads), 0, ctx->stream>>>(
        items, c + kCntAggItems, ctx->d_part_keys, ctx->d_part_counts, ctx->d_part_pairs, ctx->table_vbits,
        c + kCntOutPairs, ctx->d_keys, ctx->d_split_counts, c + kCntSplitPairs, ctx->split_cap, c + kCntFlushed);
    GW_HIP(ctx, hipGetLastError());
    return GW_OK;
}

int gw_partlog_finish(gw_ctx* ctx, const unsigned long long* cnt, bool* done, double* ms) {
    *done = false;
    *ms = 0;
    if (cnt[kCntPartOverflow]) return GW_OK;  // a region overflowed: discard
    const uint64_t split = cnt[kCntSplitPairs];
    if (split > ctx->split_cap) return gw_fail(ctx, GW_ESTATE, "gw_run: split pairs exceed the key log");
    uint64_t sums = 0;
    if (split) {
        // Flushed sources: sum their partial pairs exactly by sort-reduce.
        int rc;
        if ((rc = ensure_runs(ctx, split))) return rc;
        const int end_bit = static_cast<int>(64 - ctx->table_cbits);
        uint32_t* cnt_in = ctx->d_split_counts;
        uint32_t* cnt_sorted = ctx->d_split_counts + ctx->split_cap;
        size_t need_sort = 0, need_red = 0;
        GW_HIP(ctx, rocprim::radix_sort_pairs(nullptr, need_sort, ctx->d_keys, ctx->d_keys_alt, cnt_in,
                                                       cnt_sorted, split, 0, end_bit, ctx->stream));
        GW_HIP(ctx, rocprim::reduce_by_key(nullptr, need_red, ctx->d_keys_alt, cnt_sorted, split, ctx->d_runs, ctx->d_run_counts, ctx->d_counters + kCntSplitRuns, rocprim::plus<uint32_t>(), rocprim::equal_to<unsigned long long>(), ctx->stream));
        if ((rc = ensure_sort_temp(ctx, std::max(need_sort, need_red)))) return rc;
        GW_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
        size_t tb = ctx->sort_temp_bytes;
        GW_HIP(ctx, rocprim::radix_sort_pairs(ctx->d_sort_temp, tb, ctx->d_keys, ctx->d_keys_alt, cnt_in,
                                                       cnt_sorted, split, 0, end_bit, ctx->stream));
        tb = ctx->sort_temp_bytes;
        GW_HIP(ctx, rocprim::reduce_by_key(ctx->d_sort_temp, tb, ctx->d_keys_alt, cnt_sorted, split, ctx->d_runs, ctx->d_run_counts, ctx->d_counters + kCntSplitRuns, rocprim::plus<uint32_t>(), rocprim::equal_to<unsigned long long>(), ctx->stream));
        GW_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
        unsigned long long r = 0;
        GW_HIP(ctx, hipMemcpyAsync(&r, ctx->d_counters + kCntSplitRuns, 8, hipMemcpyDeviceToHost, ctx->stream));
        GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        float t = 0.f;
        GW_HIP(ctx, hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]));
        *ms = t;
        sums = r;
    }
    ctx->part_pairs = cnt[kCntOutPairs];
    ctx->split_runs = sums;
    ctx->n_runs = ctx->part_pairs + sums;
    ctx->part_flushed = cnt[kCntFlushed];
    *done = true;
    return GW_OK;
}

// Packs the partitioned result -- pairs at the start of each source's region
// plus the summed split pairs -- into (key, count) arrays sorted by key.
static int partitioned_sorted(gw_ctx* ctx, unsigned long long** keys_out, uint32_t** counts_out, void** mem,
                              uint64_t* mem_bytes) {
    const uint64_t n = ctx->n_runs;
    const uint64_t bytes = n * 24;  // two key arrays, two count arrays
    int rc = gw_alloc(ctx, mem, bytes, "pair read-out");
    if (rc) return rc;
    *mem_bytes = bytes;
    unsigned long long* k1 = static_cast<unsigned long long*>(*mem);
    unsigned long long* k2 = k1 + n;
    uint32_t* c1 = reinterpret_cast<uint32_t*>(k2 + n);
    uint32_t* c2 = c1 + n;
    GW_HIP(ctx, hipMemsetAsync(ctx->d_counters + kCntCompact, 0, 8, ctx->stream));
    if (ctx->part_pairs) {
        gather_regions_kernel<<<dim3(ctx->part_fields), dim3(256), 0, ctx->stream>>>(
            ctx->d_part_keys, ctx->d_part_counts, ctx->d_part_pairs, ctx->d_part_off, ctx->table_vbits, k1, c1,
            ctx->d_counters + kCntCompact);
        GW_HIP(ctx, hipGetLastError());
    }
    if (ctx->split_runs) {
        GW_HIP(ctx, hipMemcpyAsync(k1 + ctx->part_pairs, ctx->d_runs, ctx->split_runs * 8, hipMemcpyDeviceToDevice,
                                   ctx->stream));
        GW_HIP(ctx, hipMemcpyAsync(c1 + ctx->part_pairs, ctx->d_run_counts, ctx->split_runs * 4,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    }
    const int end_bit = static_cast<int>(64 - ctx->table_cbits);
    size_t need = 0;
    GW_HIP(ctx, rocprim::radix_sort_pairs(nullptr, need, k1, k2, c1, c2, n, 0, end_bit, ctx->stream));
    if ((rc = ensure_sort_temp(ctx, need))) return rc;
    size_t tb = ctx->sort_temp_bytes;
    GW_HIP(ctx, rocprim::radix_sort_pairs(ctx->d_sort_temp, tb, k1, k2, c1, c2, n, 0, end_bit, ctx->stream));
    *keys_out = k2;
    *counts_out = c2;
    return GW_OK;
}

// The last run's (source, vertex) counts on the device as keys
// (source << vbits) | vertex sorted ascending, with their counts.  Owned
// temporaries are released by release_pairs.  count_only: just n.
struct SortedPairs {
    const unsigned long long* keys = nullptr;
    const uint32_t* counts = nullptr;
    uint64_t n = 0;
    void* mem = nullptr;
    uint64_t mem_bytes = 0;
};

static void release_pairs(gw_ctx* ctx, SortedPairs& sp) {
    gw_free(ctx, sp.mem, sp.mem_bytes);
    sp = SortedPairs{};
}

static int sorted_pairs(gw_ctx* ctx, SortedPairs& sp, bool count_only) {
    sp = SortedPairs{};
    if (ctx->src_method == GW_SRC_GLOBAL) {
        sp.keys = ctx->d_runs;
        sp.counts = ctx->d_run_counts;
        sp.n = ctx->n_runs;
        return GW_OK;
    }
    if (ctx->src_method == GW_SRC_PARTITIONED) {
        sp.n = ctx->n_runs;
        if (count_only || sp.n == 0) return GW_OK;
        unsigned long long* k = nullptr;
        uint32_t* c = nullptr;
        int rc = partitioned_sorted(ctx, &k, &c, &sp.mem, &sp.mem_bytes);
        if (rc) {
            release_pairs(ctx, sp);
            return rc;
        }
        sp.keys = k;
        sp.counts = c;
        return GW_OK;
    }
    if (!ctx->d_table) return gw_fail(ctx, GW_ESTATE, "no run with GW_REC_SOURCE_COUNT");
    // hash table: compact the occupied slots, then sort by key (keys are
    // unique per slot).  Sorted from bit 0: rocPRIM's onesweep sort of 64-bit
    // keys with begin_bit > 0 returns a non-permutation on gfx950 with ROCm
    // 7.2 (measured), so it is avoided.
    const uint64_t slots = ctx->table_slots;
    const int end_bit = static_cast<int>(64 - ctx->table_cbits);
    size_t temp_bytes = 0;
    {
        unsigned long long* nk = nullptr;
        uint32_t* nc = nullptr;
        GW_HIP(ctx, rocprim::radix_sort_pairs(nullptr, temp_bytes, nk, nk, nc, nc, slots, 0, end_bit,
                                                       ctx->stream));
    }
    sp.mem_bytes = slots * 24 + temp_bytes;
    int rc = gw_alloc(ctx, &sp.mem, sp.mem_bytes, "table read-out");
    if (rc) {
        sp.mem_bytes = 0;
        return rc;
    }
    unsigned long long* ka = static_cast<unsigned long long*>(sp.mem);
    unsigned long long* kb = ka + slots;
    uint32_t* ca = reinterpret_cast<uint32_t*>(kb + slots);
    uint32_t* cb = ca + slots;
    void* temp = cb + slots;
    hipError_t e = hipMemsetAsync(ctx->d_counters + kCntCompact, 0, 8, ctx->stream);
    if (e == hipSuccess) {
        compact_kernel<<<dim3(gw_grid(slots, 256)), dim3(256), 0, ctx->stream>>>(
            ctx->d_table, slots, ctx->table_wide ? 1u : 0u, ctx->table_cbits, ka, ca, ctx->d_counters + kCntCompact);
        e = hipGetLastError();
    }
    unsigned long long n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, ctx->d_counters + kCntCompact, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && !count_only && n)
        e = rocprim::radix_sort_pairs(temp, temp_bytes, ka, kb, ca, cb, n, 0, end_bit, ctx->stream);
    if (e != hipSuccess) {
        release_pairs(ctx, sp);
        return gw_hip_check(ctx, e, "table read-out");
    }
    sp.keys = kb;
    sp.counts = cb;
    sp.n = n;
    return GW_OK;
}

extern "C" int gw_source_counts_read(gw_ctx* ctx, uint32_t* source, uint32_t* vertex, uint32_t* count, uint64_t cap,
                                     uint64_t* n_out) {
    if (!ctx || !n_out) return gw_fail(ctx, GW_EINVAL, "gw_source_counts_read: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    const bool size_only = !source && !vertex && !count;
    SortedPairs sp;
    int rc = sorted_pairs(ctx, sp, size_only);
    if (rc) return rc;
    const uint64_t n = sp.n;
    *n_out = n;
    if (size_only || n == 0) {
        release_pairs(ctx, sp);
        return GW_OK;
    }
    if (!source || !vertex || !count || cap < n) {
        release_pairs(ctx, sp);
        return gw_fail(ctx, GW_EINVAL, "gw_source_counts_read: output arrays missing or smaller than the pair count");
    }
    uint32_t* d_out = nullptr;
    if ((rc = gw_alloc(ctx, (void**)&d_out, n * 12, "pairs unpack"))) {
        release_pairs(ctx, sp);
        return rc;
    }
    unpack_runs_kernel<<<dim3(gw_grid(n, 256)), dim3(256), 0, ctx->stream>>>(
        sp.keys, sp.counts, n, ctx->table_vbits, d_out, d_out + n, d_out + 2 * n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(source, d_out, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(vertex, d_out + n, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(count, d_out + 2 * n, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    gw_free(ctx, d_out, n * 12);
    release_pairs(ctx, sp);
    if (e != hipSuccess) return gw_hip_check(ctx, e, "pairs read-out");
    return GW_OK;
}

// Per-source top-k (DiscreteDistribution::getTop, discretedistribution.hpp:72-95):
// the k most-visited vertices of each source, count descending; ties go to
// the smaller vertex id (the reference's heap keeps whichever came first,
// which depends on its racy insertion order).
extern "C" int gw_source_topk(gw_ctx* ctx, uint32_t k, uint32_t* source, uint32_t* vertex, uint32_t* count,
                              uint64_t cap, uint64_t* n_out) {
    if (!ctx || !n_out || k == 0 || k > kTopkMax)
        return gw_fail(ctx, GW_EINVAL, "gw_source_topk: bad arguments (1 <= k <= 1024)");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    SortedPairs sp;
    int rc = sorted_pairs(ctx, sp, false);
    if (rc) return rc;
    const uint64_t n = sp.n;
    if (n == 0) {
        *n_out = 0;
        release_pairs(ctx, sp);
        return GW_OK;
    }
    // segments: one per source, from a run-length encode of the key's source part
    const uint32_t vbits = ctx->table_vbits;
    uint64_t* d_src = nullptr;  // per pair: source id
    void* temp = nullptr;
    size_t temp_bytes = 0, t_rle = 0, t_scan = 0;
    const uint64_t bytes = (4 * n + 2) * 8;
    if ((rc = gw_alloc(ctx, (void**)&d_src, bytes, "top-k segments"))) {
        release_pairs(ctx, sp);
        return rc;
    }
    uint64_t* d_uniq = d_src + n;     // sources of the runs
    uint64_t* d_len = d_uniq + n;     // run lengths (n + 1 slots)
    uint64_t* d_seg = d_len + n + 1;  // segment offsets (n + 1 slots)
    auto cleanup = [&]() {
        gw_free(ctx, d_src, bytes);
        gw_free(ctx, temp, temp_bytes);
        release_pairs(ctx, sp);
    };
    source_of_kernel<<<dim3(gw_grid(n, 256)), dim3(256), 0, ctx->stream>>>(sp.keys, n, vbits, d_src);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = rocprim::run_length_encode(nullptr, t_rle, d_src, n, d_uniq, d_len, ctx->d_counters + kCntCompact, ctx->stream);
    if (e == hipSuccess) e = rocprim::exclusive_scan(nullptr, t_scan, d_len, d_seg, 0ull, n + 1, rocprim::plus<unsigned long long>(), ctx->stream);
    temp_bytes = std::max(t_rle, t_scan);
    if (e == hipSuccess && (rc = gw_alloc(ctx, &temp, temp_bytes, "top-k temp"))) {
        cleanup();
        return rc;
    }
    size_t tb = temp_bytes;
    if (e == hipSuccess)
        e = rocprim::run_length_encode(temp, tb, d_src, n, d_uniq, d_len, ctx->d_counters + kCntCompact, ctx->stream);
    unsigned long long nseg = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nseg, ctx->d_counters + kCntCompact, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    // lengths -> offsets: a zero after the last run makes the scan's last slot n
    if (e == hipSuccess) e = hipMemsetAsync(d_len + nseg, 0, 8, ctx->stream);
    tb = temp_bytes;
    if (e == hipSuccess) e = rocprim::exclusive_scan(temp, tb, d_len, d_seg, 0ull, nseg + 1, rocprim::plus<unsigned long long>(), ctx->stream);
    if (e != hipSuccess) {
        cleanup();
        return gw_hip_check(ctx, e, "top-k segments");
    }
    // top-k of every segment into [nseg x k] slots
    const uint64_t slots = nseg * k;
    uint32_t* d_top = nullptr;
    if ((rc = gw_alloc(ctx, (void**)&d_top, slots * 12 + nseg * 4, "top-k output"))) {
        cleanup();
        return rc;
    }
    topk_kernel<<<dim3(static_cast<uint32_t>(nseg)), dim3(kTopkThreads), 0, ctx->stream>>>(
        sp.keys, sp.counts, d_seg, vbits, k, d_top, d_top + slots, d_top + 2 * slots, d_top + 3 * slots);
    e = hipGetLastError();
    std::vector<uint32_t> top(slots * 3 + nseg);
    if (e == hipSuccess) e = hipMemcpyAsync(top.data(), d_top, (slots * 3 + nseg) * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    gw_free(ctx, d_top, slots * 12 + nseg * 4);
    cleanup();
    if (e != hipSuccess) return gw_hip_check(ctx, e, "top-k");
    uint64_t total = 0;
    for (uint64_t g = 0; g < nseg; ++g) total += top[slots * 3 + g];
    *n_out = total;
    if (!source && !vertex && !count) return GW_OK;
    if (!source || !vertex || !count || cap < total)
        return gw_fail(ctx, GW_EINVAL, "gw_source_topk: output arrays missing or smaller than the result");
    uint64_t o = 0;
    for (uint64_t g = 0; g < nseg; ++g)
        for (uint32_t r = 0; r < top[slots * 3 + g]; ++r, ++o) {
            source[o] = top[g * k + r];
            vertex[o] = top[slots + g * k + r];
            count[o] = top[2 * slots + g * k + r];
        }
    return GW_OK;
}
