// gw_meet.hip -- SimRank meeting counts on the device.
//
// The reference's simrank app (apps/simrank.cpp) places R walks from a and R
// from b, walk i starting at hop i*L with the restart-with-joint program
// (:30-59), records each walk's vertex per hop in updateInfo (:60-69), and
// computeResult (:71-84) adds c^l / R^2 for every walk pair (i, j) whose paths
// first coincide at offset l.  Here the walks of many pairs run in one device
// run with the per-walk event log, and one workgroup per pair counts the first
// meetings per offset: counts[k][l].  SimRank(a_k, b_k) = sum_l counts[k][l] c^l / R^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "gw_device.h"
#include "gw_internal.h"

namespace {

constexpr uint32_t kMeetThreads = 256;
constexpr uint32_t kMaxPath = 31;        // visited offsets per walk (L - 1), registers per lane
constexpr uint32_t kTileWords = 12288;   // LDS tile of b-side paths (48 KiB)

__global__ void init_pairs_kernel(gw_walk* __restrict__ w, const uint32_t* __restrict__ a,
                                  const uint32_t* __restrict__ b, uint64_t nw, uint32_t R, uint32_t L) {
    const uint64_t id = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (id >= nw) return;
    const uint64_t k = id / (2ull * R);
    const uint32_t r = static_cast<uint32_t>(id - k * 2ull * R);
    const uint32_t v = r < R ? a[k] : b[k];
    const uint32_t i = r < R ? r : r - R;
    w[id] = gw_walk{static_cast<uint32_t>(id), v, v, i * L};
}

// One workgroup per pair k, specialised on the path length P = L - 1.  The
// b-side paths are staged in LDS tiles; each lane holds one a-side path in
// registers and scans the tile's paths (the same path for every lane, an LDS
// broadcast), comparing all P offsets at once into a bit mask; the first set
// bit is the first meeting.  Meetings are rare, so they are counted with LDS
// atomics.  Offsets never visited hold no vertex and never meet (the
// reference's 0xffffffff fill).
template <uint32_t P>
__global__ __launch_bounds__(kMeetThreads) void meeting_kernel(const unsigned long long* __restrict__ ev,
                                                              uint32_t cap, uint32_t R,
                                                              const uint32_t* __restrict__ visits,
                                                              uint32_t* __restrict__ counts, uint32_t L) {
    constexpr uint32_t kStride = (P + 3) / 4 * 4;  // a b-side path padded to whole uint4s
    constexpr uint32_t kTile = kTileWords / kStride;
    __shared__ __attribute__((aligned(16))) uint32_t s_b[kTileWords];
    __shared__ uint32_t s_cnt[kMaxPath + 1];
    const uint32_t k = blockIdx.x;
    const uint64_t abase = 2ull * k * R, bbase = abase + R;
    if (threadIdx.x <= kMaxPath) s_cnt[threadIdx.x] = 0;
    for (uint32_t t0 = 0; t0 < R; t0 += kTile) {
        const uint32_t tn = std::min(kTile, R - t0);
        __syncthreads();  // the previous tile is consumed
        for (uint32_t x = threadIdx.x; x < tn * kStride; x += blockDim.x) {
            const uint32_t j = x / kStride, l = x - j * kStride;
            const uint64_t id = bbase + t0 + j;
            s_b[x] = l < P && l < visits[id] ? static_cast<uint32_t>(ev[id * cap + l]) : 0xFFFFFFFFu;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < R; i += blockDim.x) {
            const uint64_t id = abase + i;
            const uint32_t nv = visits[id];
            uint32_t a[kStride];
#pragma unroll
            for (uint32_t l = 0; l < kStride; ++l)
                a[l] = (l < P && l < nv) ? static_cast<uint32_t>(ev[id * cap + l]) : 0xFFFFFFFEu;  // never equal to a b fill
            for (uint32_t j = 0; j < tn; ++j) {
                const uint4* bp = reinterpret_cast<const uint4*>(s_b + j * kStride);
                uint32_t mask = 0;
#pragma unroll
                for (uint32_t q = 0; q < kStride / 4; ++q) {
                    const uint4 bq = bp[q];
                    mask |= (a[4 * q] == bq.x ? 1u : 0u) << (4 * q);
                    mask |= (a[4 * q + 1] == bq.y ? 1u : 0u) << (4 * q + 1);
                    mask |= (a[4 * q + 2] == bq.z ? 1u : 0u) << (4 * q + 2);
                    mask |= (a[4 * q + 3] == bq.w ? 1u : 0u) << (4 * q + 3);
                }
                if (mask) atomicAdd(&s_cnt[__ffs(mask) - 1], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < L; l += blockDim.x) counts[static_cast<uint64_t>(k) * L + l] = l < P ? s_cnt[l] : 0u;
}

// Sparse join for R <= kSparseR: only the (i, j, l) where the two paths share
// a vertex are visited.  Per offset l the a-side walks are chained by vertex:
// an open-addressing LDS table maps each distinct vertex to a slot whose head
// starts the chain of the a-side walks standing there.  Every b-side walk
// finds its vertex's slot (a few probes) and runs down that chain, all of
// whose entries meet it, test-and-setting bit (i, j) of an R x R LDS bitset
// of pairs already met; offsets are taken in increasing order, so a pair
// counts at its first meeting only.  Work per offset is R + meetings instead
// of R^2.  Each thread holds one a-side and one b-side path in registers,
// loaded up front so every load is in flight at once.
constexpr uint32_t kSparseR = 1024;
constexpr uint32_t kSparseThreads = 1024;
constexpr uint32_t kSlots = 2048;  // > 2 x the distinct vertices of one offset (<= kSparseR)
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t vertex_hash(uint32_t v) { return (v * 0x9E3779B1u) >> 21; }

// The slot of v in the offset's table, claimed if v is new.  At most kSparseR
// distinct vertices go into kSlots slots, so a free or matching slot exists.
__device__ __forceinline__ uint32_t slot_insert(uint32_t* s_key, uint32_t v) {
    uint32_t h = vertex_hash(v);
    for (uint32_t probe = 0; probe < kSlots; ++probe) {
        const uint32_t old = atomicCAS(&s_key[h], kNone, v);
        if (old == kNone || old == v) return h;
        h = (h + 1) & (kSlots - 1);
    }
    return kNone;
}

template <uint32_t P>
__global__ __launch_bounds__(kSparseThreads) void meeting_sparse_kernel(const unsigned long long* __restrict__ ev,
                                                                       uint32_t cap, uint32_t R,
                                                                       const uint32_t* __restrict__ visits,
                                                                       uint32_t* __restrict__ counts, uint32_t L) {
    __shared__ uint32_t s_bits[kSparseR * kSparseR / 32];  // 128 KiB
    __shared__ uint32_t s_key[kSlots];
    __shared__ uint32_t s_head[kSlots];
    __shared__ uint32_t s_next[kSparseR];
    __shared__ uint32_t s_cnt[kMaxPath + 1];
    const uint32_t k = blockIdx.x, tid = threadIdx.x, lane = tid & 63u;
    const uint64_t abase = 2ull * k * R, bbase = abase + R;
    const uint32_t words = (R * R + 31) / 32;
    for (uint32_t x = tid; x < words; x += blockDim.x) s_bits[x] = 0;
    if (tid <= kMaxPath) s_cnt[tid] = 0;
    // this thread's a-side and b-side walk (R <= blockDim.x); offsets never
    // visited hold kNone and never meet
    const bool has = tid < R;
    const uint32_t na = has ? visits[abase + tid] : 0u, nb = has ? visits[bbase + tid] : 0u;
    uint32_t pa[P], pb[P];
#pragma unroll
    for (uint32_t l = 0; l < P; ++l) {
        pa[l] = l < na ? static_cast<uint32_t>(ev[(abase + tid) * cap + l]) : kNone;
        pb[l] = l < nb ? static_cast<uint32_t>(ev[(bbase + tid) * cap + l]) : kNone;
    }
#pragma unroll
    for (uint32_t l = 0; l < P; ++l) {
        __syncthreads();  // the previous offset's chains are no longer read (and the bitset is clear)
        for (uint32_t x = tid; x < kSlots; x += blockDim.x) {
            s_key[x] = kNone;
            s_head[x] = kNone;
        }
        __syncthreads();
        // Walks standing on their start vertex (restarts, and every offset of
        // a start with no in-links) link up among themselves in lane order and
        // the group takes the head with one exchange.
        const uint32_t v = has ? pa[l] : kNone;
        const bool at_start = v != kNone && v == pa[0];
        const uint64_t grp = __ballot(at_start);
        if (grp) {
            const uint32_t first = __ffsll(static_cast<unsigned long long>(grp)) - 1;
            uint32_t old = kNone;
            if (lane == first) {
                const uint32_t h = slot_insert(s_key, v);
                if (h != kNone) old = atomicExch(&s_head[h], tid);
            }
            old = __shfl(old, first, 64);
            if (at_start) {
                const uint64_t above = grp & ~((2ull << lane) - 1);  // members in higher lanes
                s_next[tid] = above ? (tid - lane) + (__ffsll(static_cast<unsigned long long>(above)) - 1) : old;
            }
        }
        if (v != kNone && !at_start) {
            const uint32_t h = slot_insert(s_key, v);
            if (h != kNone) s_next[tid] = atomicExch(&s_head[h], tid);
        }
        __syncthreads();
        uint32_t met = 0;
        const uint32_t vb = has ? pb[l] : kNone;
        if (vb != kNone) {
            uint32_t h = vertex_hash(vb);
            for (uint32_t probe = 0; probe < kSlots; ++probe) {
                const uint32_t key = s_key[h];
                if (key == kNone) break;
                if (key == vb) {
                    for (uint32_t x = s_head[h]; x != kNone; x = s_next[x]) {
                        const uint32_t idx = x * R + tid;
                        const uint32_t bit = 1u << (idx & 31u);
                        met += (atomicOr(&s_bits[idx >> 5], bit) & bit) ? 0u : 1u;
                    }
                    break;
                }
                h = (h + 1) & (kSlots - 1);
            }
        }
        met = gw::wave_sum(met);
        if (lane == 0 && met) atomicAdd(&s_cnt[l], met);
    }
    __syncthreads();
    for (uint32_t l = tid; l < L; l += blockDim.x) counts[static_cast<uint64_t>(k) * L + l] = l < P ? s_cnt[l] : 0u;
}

template <uint32_t P>
hipError_t launch_meeting(uint32_t npairs, hipStream_t st, const unsigned long long* ev, uint32_t cap, uint32_t R,
                          const uint32_t* visits, uint32_t* counts, uint32_t L) {
    if (R <= kSparseR)
        meeting_sparse_kernel<P><<<dim3(npairs), dim3(kSparseThreads), 0, st>>>(ev, cap, R, visits, counts, L);
    else
        meeting_kernel<P><<<dim3(npairs), dim3(kMeetThreads), 0, st>>>(ev, cap, R, visits, counts, L);
    return hipGetLastError();
}

template <uint32_t... Ps>
hipError_t dispatch_meeting(uint32_t P, std::integer_sequence<uint32_t, Ps...>, uint32_t npairs, hipStream_t st,
                            const unsigned long long* ev, uint32_t cap, uint32_t R, const uint32_t* visits,
                            uint32_t* counts, uint32_t L) {
    hipError_t e = hipErrorInvalidValue;
    ((P == Ps + 1 ? (e = launch_meeting<Ps + 1>(npairs, st, ev, cap, R, visits, counts, L), 0) : 0), ...);
    return e;
}

}  // namespace

extern "C" {

int gw_walks_init_pairs(gw_ctx* ctx, const uint32_t* a, const uint32_t* b, uint32_t npairs, uint32_t R, uint32_t L) {
    if (!ctx || !a || !b || npairs == 0 || R == 0 || L == 0)
        return gw_fail(ctx, GW_EINVAL, "gw_walks_init_pairs: bad arguments");
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_walks_init_pairs: no graph");
    const uint64_t nw = 2ull * npairs * R;
    if (nw >= 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_walks_init_pairs: walk ids are 32-bit");
    if (static_cast<uint64_t>(R) * L >= 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_walks_init_pairs: R * L overflows the hop field");
    uint32_t src_max = 0;
    for (uint32_t k = 0; k < npairs; ++k) {
        if (a[k] >= ctx->n || b[k] >= ctx->n) return gw_fail(ctx, GW_EINVAL, "gw_walks_init_pairs: vertex out of range");
        src_max = std::max(src_max, std::max(a[k], b[k]));
    }
    GW_HIP(ctx, hipSetDevice(ctx->device));
    int rc = gw_reserve_walks(ctx, nw);
    if (rc) return rc;
    uint32_t* d_ab = nullptr;
    if ((rc = gw_alloc(ctx, (void**)&d_ab, npairs * 8ull, "pair table"))) return rc;
    hipError_t e = hipMemcpyAsync(d_ab, a, npairs * 4ull, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ab + npairs, b, npairs * 4ull, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        init_pairs_kernel<<<dim3(static_cast<uint32_t>((nw + 255) / 256)), dim3(256), 0, ctx->stream>>>(
            ctx->d_walks, d_ab, d_ab + npairs, nw, R, L);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    gw_free(ctx, d_ab, npairs * 8ull);
    if (e != hipSuccess) return gw_hip_check(ctx, e, "init_pairs_kernel");
    ctx->walk_src_max = src_max;
    ctx->walk_v_max = src_max;
    ctx->field_walks.clear();
    ctx->field_vertex.clear();
    ctx->part_per_walk = 0;
    return GW_OK;
}

int gw_meeting_counts(gw_ctx* ctx, uint32_t npairs, uint32_t R, uint32_t L, uint32_t* counts) {
    if (!ctx || !counts || npairs == 0 || R == 0 || L < 2) return gw_fail(ctx, GW_EINVAL, "gw_meeting_counts: bad arguments");
    const uint32_t P = L - 1;
    if (P > kMaxPath) return gw_fail(ctx, GW_EINVAL, "gw_meeting_counts: L - 1 must be at most 31");
    if (ctx->pool || ctx->nwalks != 2ull * npairs * R)
        return gw_fail(ctx, GW_ESTATE, "gw_meeting_counts: the loaded walks are not npairs x 2R pair walks");
    if (!ctx->d_events || !ctx->d_visits || ctx->events_cap < P)
        return gw_fail(ctx, GW_ESTATE, "gw_meeting_counts: needs a run with GW_REC_EVENTS and event_cap >= L - 1");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    const uint64_t need = static_cast<uint64_t>(npairs) * L * 4;
    if (need > ctx->meet_bytes) {  // kept for the next call
        gw_free(ctx, ctx->d_meet, ctx->meet_bytes);
        ctx->d_meet = nullptr;
        ctx->meet_bytes = 0;
        int rc = gw_alloc(ctx, (void**)&ctx->d_meet, need, "meeting counts");
        if (rc) return rc;
        ctx->meet_bytes = need;
    }
    uint32_t* d_counts = ctx->d_meet;
    hipError_t e = dispatch_meeting(P, std::make_integer_sequence<uint32_t, kMaxPath>{}, npairs, ctx->stream,
                                    reinterpret_cast<const unsigned long long*>(ctx->d_events), ctx->events_cap, R,
                                    ctx->d_visits, d_counts, L);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, d_counts, need, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return gw_hip_check(ctx, e, "meeting_kernel");
    return GW_OK;
}

}  // extern "C"
