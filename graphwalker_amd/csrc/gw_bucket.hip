// gw_bucket.hip -- stable counting sort of walk records by the block of their
// current vertex: the device form of moveWalk into per-block buffers plus the
// bucket gather of getCurrentWalks (src/walks/walk.hpp:93-137).
//
// Three passes over tiles of kTile records:
//   hist     per-tile histogram over blocks.  Lanes of a wavefront that hold
//            the same block are grouped with ballots (one LDS atomic per
//            distinct block per wavefront, not per record).
//   scan     exclusive scan of the block-major histogram matrix
//            hist[b * ntiles + t]: bucket b's tiles in tile order, so the
//            scatter below is stable.
//   scatter  each record's rank among equal-block records of its tile, in
//            tile order: ballot rank inside the wavefront, plus earlier
//            wavefronts' counts, plus earlier iterations' running counts.
// Records whose vertex lies outside [block_first[0], block_first[nblocks])
// fall into bucket `nblocks` and are dropped (getblock returning nblocks,
// randomwalk.hpp:74-82; the caller returns without moveWalk).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gw_device.h"
#include "gw_internal.h"

namespace gw {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kWaves = kThreads / 64;
constexpr uint32_t kPerThread = 16;
constexpr uint32_t kTile = kThreads * kPerThread;
constexpr uint32_t kMaxBlocks = 2048;
constexpr uint32_t kScanThreads = 1024;

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint32_t find_block(const uint32_t* bf, uint32_t nb, uint32_t v) {
    if (v < bf[0] || v >= bf[nb]) return nb;
    uint32_t lo = 0, hi = nb;  // bf[lo] <= v < bf[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bf[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

// Group the lanes of a wavefront by key.  For each valid lane returns its rank
// among the lanes holding the same key and whether it is the group's first
// lane; `count` receives the group size (valid in every lane of the group).
__device__ __forceinline__ void wave_group(uint32_t key, bool valid, uint32_t& rank, uint32_t& count, bool& first) {
    uint64_t todo = __ballot(valid);
    rank = 0;
    count = 0;
    first = false;
    while (todo) {
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
        const uint32_t k = __shfl(key, leader, 64);
        const uint64_t eq = __ballot(valid && key == k);
        if (valid && key == k) {
            rank = lane_rank(eq);
            count = __popcll(eq);
            first = (rank == 0);
        }
        todo &= ~eq;
    }
}

// minhop (optional): per-block minimum hop of the records, the minstep table
// of the reference's block choice (walk.hpp:187-202), kept with atomicMin.
__global__ __launch_bounds__(kThreads) void hist_kernel(const gw_walk* __restrict__ in, uint64_t n,
                                                        const uint32_t* __restrict__ block_first, uint32_t nb,
                                                        uint32_t ntiles, uint32_t* __restrict__ hist,
                                                        uint32_t* __restrict__ minhop) {
    __shared__ uint32_t s_bf[kMaxBlocks + 1];
    __shared__ uint32_t s_cnt[kMaxBlocks + 1];
    __shared__ uint32_t s_min[kMaxBlocks + 1];
    for (uint32_t i = threadIdx.x; i <= nb; i += kThreads) {
        s_bf[i] = block_first[i];
        s_cnt[i] = 0;
        s_min[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTile;
    for (uint32_t it = 0; it < kPerThread; ++it) {
        const uint64_t i = base + it * kThreads + threadIdx.x;
        const bool valid = i < n;
        const gw_walk w = valid ? in[i] : gw_walk{0, 0, 0, 0};
        const uint32_t b = valid ? find_block(s_bf, nb, w.vertex) : 0u;
        uint32_t rank, count;
        bool first;
        wave_group(b, valid, rank, count, first);
        if (valid && first) atomicAdd(&s_cnt[b], count);
        if (minhop && valid && w.hop < s_min[b]) atomicMin(&s_min[b], w.hop);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b <= nb; b += kThreads) {
        hist[static_cast<uint64_t>(b) * ntiles + blockIdx.x] = s_cnt[b];
        if (minhop && s_min[b] != 0xFFFFFFFFu) atomicMin(&minhop[b], s_min[b]);
    }
}

// Exclusive scan, phase 1: each workgroup scans its chunk of kScanThreads*4
// entries in place and writes the chunk total.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_wave, uint32_t& total) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64, nw = blockDim.x / 64;
    const uint32_t inc = wave_incl_scan(x);
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t w = threadIdx.x < nw ? s_wave[threadIdx.x] : 0u;
        const uint32_t wi = wave_incl_scan(w);
        if (threadIdx.x < nw) s_wave[threadIdx.x] = wi - w;
        if (threadIdx.x == nw - 1) s_wave[32] = wi;
    }
    __syncthreads();
    total = s_wave[32];
    const uint32_t r = s_wave[wave] + inc - x;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanThreads) void scan_chunks_kernel(uint32_t* __restrict__ a, uint64_t n,
                                                                   uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_wave[33];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanThreads * 4 + threadIdx.x * 4ull;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? a[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(s, s_wave, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) a[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Phase 2: one workgroup scans the chunk totals (up to kScanThreads*4 chunks);
// phase 3 adds them back.
__global__ __launch_bounds__(kScanThreads) void scan_sums_kernel(uint32_t* __restrict__ sums, uint32_t nchunks,
                                                                 unsigned long long* __restrict__ grand) {
    __shared__ uint32_t s_wave[33];
    uint32_t v[4], s = 0;
    const uint32_t base = threadIdx.x * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < nchunks ? sums[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(s, s_wave, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < nchunks) sums[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) *grand = total;
}

__global__ void scan_add_kernel(uint32_t* __restrict__ a, uint64_t n, const uint32_t* __restrict__ sums) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) a[i] += sums[i / (kScanThreads * 4)];
}

__global__ __launch_bounds__(kThreads) void scatter_kernel(const gw_walk* __restrict__ in, uint64_t n,
                                                           const uint32_t* __restrict__ block_first, uint32_t nb,
                                                           uint32_t ntiles, const uint32_t* __restrict__ offs,
                                                           gw_walk* __restrict__ out) {
    __shared__ uint32_t s_bf[kMaxBlocks + 1];
    __shared__ uint32_t s_run[kMaxBlocks + 1];
    __shared__ uint32_t s_wcnt[kWaves][kMaxBlocks + 1];
    const uint32_t wave = threadIdx.x / 64;
    for (uint32_t i = threadIdx.x; i <= nb; i += kThreads) {
        s_bf[i] = block_first[i];
        s_run[i] = offs[static_cast<uint64_t>(i) * ntiles + blockIdx.x];
        for (uint32_t w = 0; w < kWaves; ++w) s_wcnt[w][i] = 0;
    }
    __syncthreads();
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTile;
    for (uint32_t it = 0; it < kPerThread; ++it) {
        const uint64_t i = base + it * kThreads + threadIdx.x;
        const bool valid = i < n;
        gw_walk w{0, 0, 0, 0};
        if (valid) w = in[i];
        const uint32_t b = valid ? find_block(s_bf, nb, w.vertex) : 0u;
        uint32_t rank, count;
        bool first;
        wave_group(b, valid, rank, count, first);
        if (valid && first) s_wcnt[wave][b] = count;
        __syncthreads();
        if (valid && b < nb) {
            uint32_t pos = s_run[b] + rank;
            for (uint32_t ww = 0; ww < wave; ++ww) pos += s_wcnt[ww][b];
            out[pos] = w;
        }
        __syncthreads();
        // every group leader moves its block's running offset past its group
        // and clears its wave count for the next iteration
        if (valid && first) {
            atomicAdd(&s_run[b], count);
            s_wcnt[wave][b] = 0;
        }
        __syncthreads();
    }
}

__global__ void gather_starts_kernel(const uint32_t* __restrict__ offs, uint32_t ntiles, uint32_t nb,
                                     uint32_t* __restrict__ starts) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= nb) starts[b] = offs[static_cast<uint64_t>(b) * ntiles];
}

}  // namespace
}  // namespace gw

int gw_bucket_sort(gw_ctx* ctx, const gw_walk* in, uint64_t n, gw_walk* out, std::vector<uint64_t>& bucket_off,
                   std::vector<uint32_t>* minhop) {
    using namespace gw;
    const uint32_t nb = ctx->nblocks;
    if (nb == 0 || nb > kMaxBlocks) return gw_fail(ctx, GW_EINVAL, "bucketing supports 1..2048 blocks");
    const uint32_t ntiles = static_cast<uint32_t>(std::max<uint64_t>(1, (n + kTile - 1) / kTile));
    const uint64_t entries = static_cast<uint64_t>(nb + 1) * ntiles;
    const uint64_t nchunks = (entries + kScanThreads * 4 - 1) / (kScanThreads * 4);
    if (nchunks > kScanThreads * 4) return gw_fail(ctx, GW_EINVAL, "bucketing: too many records for the two-level scan");
    const uint64_t need = entries + nchunks + nb + 1;
    if (need > ctx->tile_hist_cap) {
        gw_free(ctx, ctx->d_tile_hist, ctx->tile_hist_cap * 4);
        ctx->d_tile_hist = nullptr;
        ctx->tile_hist_cap = 0;
        int rc = gw_alloc(ctx, (void**)&ctx->d_tile_hist, need * 4, "bucket histogram");
        if (rc) return rc;
        ctx->tile_hist_cap = need;
    }
    uint32_t* hist = ctx->d_tile_hist;
    uint32_t* sums = hist + entries;
    uint32_t* d_min = nullptr;
    if (minhop) {
        if (ctx->block_minhop_cap < nb + 1) {
            gw_free(ctx, ctx->d_block_minhop, ctx->block_minhop_cap * 4ull);
            ctx->d_block_minhop = nullptr;
            ctx->block_minhop_cap = 0;
            int rc = gw_alloc(ctx, (void**)&ctx->d_block_minhop, (kMaxBlocks + 1) * 4ull, "block min hop");
            if (rc) return rc;
            ctx->block_minhop_cap = kMaxBlocks + 1;
        }
        d_min = ctx->d_block_minhop;
        GW_HIP(ctx, hipMemsetAsync(d_min, 0xFF, (nb + 1) * 4ull, ctx->stream));
    }
    hist_kernel<<<dim3(ntiles), dim3(kThreads), 0, ctx->stream>>>(in, n, ctx->d_block_first, nb, ntiles, hist, d_min);
    GW_HIP(ctx, hipGetLastError());
    // block-major sizes: read back the per-bucket totals before the scan
    // overwrites the histogram, by scanning and reading bucket starts.
    scan_chunks_kernel<<<dim3(static_cast<uint32_t>(nchunks)), dim3(kScanThreads), 0, ctx->stream>>>(hist, entries, sums);
    GW_HIP(ctx, hipGetLastError());
    scan_sums_kernel<<<dim3(1), dim3(kScanThreads), 0, ctx->stream>>>(sums, static_cast<uint32_t>(nchunks), ctx->d_counters + kCntBucketTotal);
    GW_HIP(ctx, hipGetLastError());
    scan_add_kernel<<<dim3(static_cast<uint32_t>((entries + 255) / 256)), dim3(256), 0, ctx->stream>>>(hist, entries, sums);
    GW_HIP(ctx, hipGetLastError());
    scatter_kernel<<<dim3(ntiles), dim3(kThreads), 0, ctx->stream>>>(in, n, ctx->d_block_first, nb, ntiles, hist, out);
    GW_HIP(ctx, hipGetLastError());
    // bucket b starts at hist[b * ntiles] (scanned); the dropped bucket nb
    // starts where the kept records end.
    uint32_t* d_starts = sums + nchunks;
    gather_starts_kernel<<<dim3((nb + 1 + 255) / 256), dim3(256), 0, ctx->stream>>>(hist, ntiles, nb, d_starts);
    GW_HIP(ctx, hipGetLastError());
    std::vector<uint32_t> starts(nb + 1);
    GW_HIP(ctx, hipMemcpyAsync(starts.data(), d_starts, (nb + 1) * 4ull, hipMemcpyDeviceToHost, ctx->stream));
    if (minhop) {
        minhop->resize(nb + 1);
        GW_HIP(ctx, hipMemcpyAsync(minhop->data(), d_min, (nb + 1) * 4ull, hipMemcpyDeviceToHost, ctx->stream));
    }
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    bucket_off.assign(starts.begin(), starts.end());
    return GW_OK;
}

extern "C" int gw_bucket(gw_ctx* ctx, uint64_t* bucket_counts) {
    if (!ctx) return GW_EINVAL;
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_bucket: no graph");
    if (ctx->pool) return gw_fail(ctx, GW_ESTATE, "gw_bucket: the walks are in a host pool");
    if (ctx->nwalks >= 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_bucket: too many walks");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<uint64_t> off;
    int rc = gw_bucket_sort(ctx, ctx->d_walks, ctx->nwalks, ctx->d_walks_alt, off);
    if (rc) return rc;
    std::swap(ctx->d_walks, ctx->d_walks_alt);
    gw_gen_clear(ctx);  // the walks are in block order now
    const uint32_t nb = ctx->nblocks;
    const uint64_t kept = off[nb];
    for (uint32_t b = 0; b < nb; ++b) {
        const uint64_t end = b + 1 < nb ? off[b + 1] : kept;
        if (bucket_counts) bucket_counts[b] = end - off[b];
    }
    ctx->nwalks = kept;
    return GW_OK;
}
