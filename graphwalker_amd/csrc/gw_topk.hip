// gw_topk.hip -- top-k vertices of the per-vertex counts on the device.
//
// The reference's get_top_vertices (src/util/toplist.hpp:86-183) streams the
// vertex-value file in windows of 2^20, quick-sorts each window's candidates
// above the running minimum and merges them into a top list.  Here the counts
// are already in HBM after a GW_REC_VERTEX_COUNT run:
//   1. MSB radix select over the u32 counts, four 8-bit digits: one streaming
//      histogram pass per digit (restricted to the prefix chosen so far) finds
//      the k-th largest count T, how many counts exceed it (g < k) and how many
//      equal-to-T vertices are still needed (r = k - g);
//   2. ordered compaction: per-chunk match counts, an exclusive scan, then each
//      chunk writes its matches at its offset in vertex order (first the g
//      vertices with count > T, then the first r vertices with count == T);
//   3. the g larger ones are radix-sorted by (count descending, vertex
//      ascending) and the ties follow with count T.
// The order is total (count descending, then vertex id ascending); the
// reference's quickSort leaves ties among equal values in no fixed order.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "gw_internal.h"

namespace {

constexpr uint32_t kHistThreads = 256;
constexpr uint32_t kChunkThreads = 256;
constexpr uint32_t kChunkIters = 16;
constexpr uint64_t kChunk = static_cast<uint64_t>(kChunkThreads) * kChunkIters;  // vertices per chunk

__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t c, uint32_t shift, uint32_t hi_mask, uint32_t prefix,
                                         uint32_t& cur, uint32_t& run) {
    if ((c & hi_mask) != prefix) return;
    const uint32_t d = (c >> shift) & 255u;
    if (d == cur) {
        ++run;
    } else {
        if (run) atomicAdd(&h[cur], run);
        cur = d;
        run = 1;
    }
}

// Histogram of digit (c >> shift) & 255 over the counts whose bits above the
// digit equal prefix.  A thread keeps a run of equal digits in registers and
// adds it to LDS when the digit changes: skewed count vectors (mostly zeros)
// then make few LDS atomics.
__global__ void digit_hist_kernel(const uint32_t* __restrict__ counts, uint64_t n, uint32_t shift, uint32_t hi_mask,
                                  uint32_t prefix, unsigned long long* __restrict__ hist) {
    __shared__ uint32_t h[256];
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
    __syncthreads();
    uint32_t cur = 0, run = 0;
    const uint64_t n4 = n / 4;
    const uint4* c4 = reinterpret_cast<const uint4*>(counts);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t t0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    for (uint64_t i = t0; i < n4; i += stride) {
        const uint4 q = c4[i];
        hist_add(h, q.x, shift, hi_mask, prefix, cur, run);
        hist_add(h, q.y, shift, hi_mask, prefix, cur, run);
        hist_add(h, q.z, shift, hi_mask, prefix, cur, run);
        hist_add(h, q.w, shift, hi_mask, prefix, cur, run);
    }
    for (uint64_t i = n4 * 4 + t0; i < n; i += stride) hist_add(h, counts[i], shift, hi_mask, prefix, cur, run);
    if (run) atomicAdd(&h[cur], run);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], static_cast<unsigned long long>(h[i]));
}

__device__ __forceinline__ bool topk_match(uint32_t c, uint32_t T, uint32_t ties) { return ties ? c == T : c > T; }

// Matches per chunk of kChunk vertices.
__global__ void chunk_count_kernel(const uint32_t* __restrict__ counts, uint64_t n, uint32_t T, uint32_t ties,
                                   uint32_t* __restrict__ chunk_cnt, uint64_t nchunks) {
    __shared__ uint32_t wsum[kChunkThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        uint32_t m = 0;
        for (uint32_t it = 0; it < kChunkIters; ++it) {
            const uint64_t v = c * kChunk + it * static_cast<uint64_t>(kChunkThreads) + threadIdx.x;
            m += (v < n && topk_match(counts[v], T, ties)) ? 1u : 0u;
        }
        for (uint32_t o = 32; o > 0; o >>= 1) m += __shfl_down(m, o, 64);
        if (lane == 0) wsum[wave] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t s = 0;
            for (uint32_t w = 0; w < kChunkThreads / 64; ++w) s += wsum[w];
            chunk_cnt[c] = s;
        }
        __syncthreads();
    }
}

// Each chunk writes its matching vertex ids, in vertex order, at
// out[chunk_off[c] ...], keeping those below limit.
__global__ void chunk_write_kernel(const uint32_t* __restrict__ counts, uint64_t n, uint32_t T, uint32_t ties,
                                   const uint64_t* __restrict__ chunk_off, uint64_t nchunks, uint64_t limit,
                                   uint32_t* __restrict__ out) {
    __shared__ uint32_t wsum[kChunkThreads / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        uint64_t base = chunk_off[c];
        if (base >= limit) continue;  // uniform across the block
        for (uint32_t it = 0; it < kChunkIters; ++it) {
            const uint64_t v = c * kChunk + it * static_cast<uint64_t>(kChunkThreads) + threadIdx.x;
            const bool f = v < n && topk_match(counts[v], T, ties);
            const uint64_t mask = __ballot(f);
            if (lane == 0) wsum[wave] = static_cast<uint32_t>(__popcll(mask));
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (uint32_t w = 0; w < kChunkThreads / 64; ++w) {
                before += w < wave ? wsum[w] : 0u;
                total += wsum[w];
            }
            const uint64_t pos = base + before + static_cast<uint32_t>(__popcll(mask & below));
            if (f && pos < limit) out[pos] = static_cast<uint32_t>(v);
            base += total;
            __syncthreads();
        }
    }
}

// key = (~count << 32) | vertex: ascending keys = count descending, vertex ascending.
__global__ void topk_keys_kernel(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ verts, uint64_t g,
                                 uint64_t* __restrict__ keys) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < g; i += stride) {
        const uint32_t v = verts[i];
        keys[i] = (static_cast<uint64_t>(~counts[v]) << 32) | v;
    }
}

__global__ void topk_unpack_kernel(const uint64_t* __restrict__ keys, uint64_t g, uint32_t T,
                                   const uint32_t* __restrict__ tie_verts, uint64_t r, uint32_t* __restrict__ vert,
                                   uint32_t* __restrict__ cnt) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < g + r; i += stride) {
        if (i < g) {
            vert[i] = static_cast<uint32_t>(keys[i]);
            cnt[i] = ~static_cast<uint32_t>(keys[i] >> 32);
        } else {
            vert[i] = tie_verts[i - g];
            cnt[i] = T;
        }
    }
}

// Device buffers of one call, released together.
struct TopkBufs {
    gw_ctx* ctx;
    std::vector<std::pair<void*, uint64_t>> held;
    template <class T>
    int get(T** p, uint64_t bytes, const char* what) {
        bytes = std::max<uint64_t>(bytes, 16);
        int rc = gw_alloc(ctx, reinterpret_cast<void**>(p), bytes, what);
        if (!rc) held.emplace_back(*p, bytes);
        return rc;
    }
    ~TopkBufs() {
        for (auto& h : held) gw_free(ctx, h.first, h.second);
    }
};

// Ordered compaction of the vertices matching (count > T) or (count == T)
// into out[0 .. limit).
int compact_ordered(gw_ctx* ctx, TopkBufs& B, uint32_t T, uint32_t ties, uint64_t limit, uint32_t* out) {
    const uint64_t n = ctx->n;
    const uint64_t nchunks = (n + kChunk - 1) / kChunk;
    uint32_t* cnt = nullptr;
    uint64_t* off = nullptr;
    int rc;
    if ((rc = B.get(&cnt, nchunks * 4, "top-k chunk counts")) || (rc = B.get(&off, nchunks * 8, "top-k chunk offsets")))
        return rc;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(nchunks, std::max(1u, ctx->cu_count) * 8ull));
    chunk_count_kernel<<<dim3(grid), dim3(kChunkThreads), 0, ctx->stream>>>(ctx->d_counts, n, T, ties, cnt, nchunks);
    GW_HIP(ctx, hipGetLastError());
    size_t need = 0;
    GW_HIP(ctx, rocprim::exclusive_scan(nullptr, need, cnt, off, 0ull, nchunks, rocprim::plus<unsigned long long>(), ctx->stream));
    void* temp = nullptr;
    if ((rc = B.get(&temp, need, "top-k scan temp"))) return rc;
    GW_HIP(ctx, rocprim::exclusive_scan(temp, need, cnt, off, 0ull, nchunks, rocprim::plus<unsigned long long>(), ctx->stream));
    chunk_write_kernel<<<dim3(grid), dim3(kChunkThreads), 0, ctx->stream>>>(ctx->d_counts, n, T, ties, off, nchunks,
                                                                           limit, out);
    GW_HIP(ctx, hipGetLastError());
    return GW_OK;
}

}  // namespace

extern "C" int gw_counts_topk(gw_ctx* ctx, uint64_t k, uint32_t* vertex, uint32_t* count, uint64_t* n_out) {
    if (!ctx || k == 0 || !vertex || !count || !n_out) return gw_fail(ctx, GW_EINVAL, "gw_counts_topk: bad arguments");
    if (!ctx->d_counts || !gw_has_graph(ctx))
        return gw_fail(ctx, GW_ESTATE, "gw_counts_topk: no per-vertex counts (run a GW_REC_VERTEX_COUNT program first)");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    const uint64_t n = ctx->n;
    k = std::min<uint64_t>(k, n);
    *n_out = 0;
    if (k == 0) return GW_OK;
    TopkBufs B{ctx, {}};
    int rc;
    unsigned long long* d_hist = nullptr;
    if ((rc = B.get(&d_hist, 256 * 8, "top-k histogram"))) return rc;
    std::vector<unsigned long long> hist(256);
    // 1. the k-th largest count T, digit by digit from the top
    uint64_t remaining = k;
    uint32_t prefix = 0;
    const uint32_t grid = static_cast<uint32_t>(
        std::max<uint64_t>(1, std::min<uint64_t>((n / 4 + kHistThreads - 1) / kHistThreads, std::max(1u, ctx->cu_count) * 4ull)));
    for (int shift = 24; shift >= 0; shift -= 8) {
        const uint32_t hi_mask = shift == 24 ? 0u : (0xFFFFFFFFu << (shift + 8));
        GW_HIP(ctx, hipMemsetAsync(d_hist, 0, 256 * 8, ctx->stream));
        digit_hist_kernel<<<dim3(grid), dim3(kHistThreads), 0, ctx->stream>>>(ctx->d_counts, n,
                                                                               static_cast<uint32_t>(shift), hi_mask,
                                                                               prefix, d_hist);
        GW_HIP(ctx, hipGetLastError());
        GW_HIP(ctx, hipMemcpyAsync(hist.data(), d_hist, 256 * 8, hipMemcpyDeviceToHost, ctx->stream));
        GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        int d = 255;
        for (; d > 0; --d) {
            if (hist[d] >= remaining) break;
            remaining -= hist[d];
        }
        prefix |= static_cast<uint32_t>(d) << shift;
    }
    const uint32_t T = prefix;
    const uint64_t g = k - remaining;  // vertices with count > T
    // 2. ordered compaction of both groups
    uint32_t *gv = nullptr, *tv = nullptr, *ov = nullptr, *oc = nullptr;
    uint64_t *keys = nullptr, *keys_out = nullptr;
    if ((rc = B.get(&gv, g * 4, "top-k larger")) || (rc = B.get(&tv, remaining * 4, "top-k ties")) ||
        (rc = B.get(&ov, k * 4, "top-k vertices")) || (rc = B.get(&oc, k * 4, "top-k counts")) ||
        (rc = B.get(&keys, g * 8, "top-k keys")) || (rc = B.get(&keys_out, g * 8, "top-k keys (sorted)")))
        return rc;
    if (g && (rc = compact_ordered(ctx, B, T, 0u, g, gv))) return rc;
    if ((rc = compact_ordered(ctx, B, T, 1u, remaining, tv))) return rc;
    // 3. sort the larger ones, append the ties
    const uint32_t tb = 256;
    if (g) {
        topk_keys_kernel<<<dim3(gw_grid(g, tb)), dim3(tb), 0, ctx->stream>>>(ctx->d_counts, gv, g, keys);
        GW_HIP(ctx, hipGetLastError());
        size_t need = 0;
        GW_HIP(ctx, rocprim::radix_sort_keys(nullptr, need, keys, keys_out, g, 0, 64, ctx->stream));
        void* temp = nullptr;
        if ((rc = B.get(&temp, need, "top-k sort temp"))) return rc;
        GW_HIP(ctx, rocprim::radix_sort_keys(temp, need, keys, keys_out, g, 0, 64, ctx->stream));
    }
    topk_unpack_kernel<<<dim3(gw_grid(k, tb)), dim3(tb), 0, ctx->stream>>>(keys_out, g, T, tv, remaining, ov, oc);
    GW_HIP(ctx, hipGetLastError());
    GW_HIP(ctx, hipMemcpyAsync(vertex, ov, k * 4, hipMemcpyDeviceToHost, ctx->stream));
    GW_HIP(ctx, hipMemcpyAsync(count, oc, k * 4, hipMemcpyDeviceToHost, ctx->stream));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *n_out = k;
    return GW_OK;
}
