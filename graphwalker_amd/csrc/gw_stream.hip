// gw_stream.hip -- out-of-core graphs: the CSR stays in pinned host memory
// and a device block cache holds the blocks a round steps.
//
// The reference engine keeps `nmblocks` blocks in memory, reads a missing
// block's beg_pos / csr slices from disk and evicts the resident block with
// the fewest walks (findSubGraph / loadSubGraph / swapOut,
// src/engine/graphwalker.hpp:168-251).  Here the cache is a set of equal
// device slots sized for the largest block; gw_step.hip's streamed rounds
// choose the blocks, copy them over PCIe on a side stream (hipMemcpyAsync
// from pinned memory, default stream priority) and step their buckets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "gw_device.h"
#include "gw_internal.h"

namespace {

constexpr uint64_t kSlotAlign = 256;

uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

uint64_t block_bytes(const gw_ctx* ctx, uint32_t b, uint64_t* off_bytes) {
    const uint32_t lo = ctx->block_first[b], hi = ctx->block_first[b + 1];
    const uint64_t ob = align_up((hi - lo + 1ull) * 8, kSlotAlign);
    if (off_bytes) *off_bytes = ob;
    return ob + (ctx->h_off[hi] - ctx->h_off[lo]) * 4;
}

// Greedy cut of the vertices into blocks of at most `target` bytes (offsets +
// columns), each at least one vertex: the rule of compute_block
// (conversions.hpp:300-364) with bytes in place of the file block size.
std::vector<uint32_t> cut_blocks(const gw_ctx* ctx, uint64_t target) {
    std::vector<uint32_t> bf{0};
    uint32_t lo = 0;
    while (lo < ctx->n) {
        uint32_t hi = lo + 1;
        // grow while the block [lo, hi + 1) still fits
        while (hi < ctx->n && (hi + 1 - lo + 1ull) * 8 + (ctx->h_off[hi + 1] - ctx->h_off[lo]) * 4 <= target) ++hi;
        bf.push_back(hi);
        lo = hi;
    }
    return bf;
}

// Independent random 4-byte reads of pinned host memory by a kernel: index
// i's address comes from Philox(i), so a lane's reads are all in flight
// together (the column reads of pages a server reads from host memory).
__global__ void host_gather_kernel(const uint32_t* __restrict__ buf, uint32_t nwords, uint64_t gathers,
                                   uint32_t* __restrict__ out) {
    const uint64_t tid = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    uint32_t acc = 0;
    for (uint64_t i = tid; i < gathers; i += stride) {
        const gw::u32x4 r = gw::philox4x32_10(static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), 0x484F5354u, 0, 3, 4);
        acc ^= buf[gw::mulhi32(r.x, nwords)];
    }
    out[tid] = acc;
}

}  // namespace

bool gw_is_pinned(const gw_ctx* ctx, const void* p, uint64_t bytes) {
    const char* c = static_cast<const char*>(p);
    for (const auto& a : ctx->pinned) {
        const char* b = static_cast<const char*>(a.first);
        if (c >= b && c + bytes <= b + a.second) return true;
    }
    return false;
}

void gw_stream_reset(gw_ctx* ctx) {
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    for (void* m : ctx->slot_mem) gw_free(ctx, m, ctx->slot_bytes);
    ctx->slot_mem.clear();
    ctx->slot_block.clear();
    ctx->block_slot.clear();
    ctx->slot_bytes = 0;
    if (ctx->h_views) (void)hipHostFree(ctx->h_views);
    ctx->h_views = nullptr;
    gw_free(ctx, ctx->d_views, 0);
    ctx->d_views = nullptr;
}

void gw_host_graph_free(gw_ctx* ctx) {
    gw_stream_reset(ctx);
    if (ctx->h_owned_off) (void)hipHostFree(ctx->h_owned_off);
    if (ctx->h_owned_col) (void)hipHostFree(ctx->h_owned_col);
    ctx->h_owned_off = ctx->h_owned_col = nullptr;
    ctx->h_off = nullptr;
    ctx->h_col = nullptr;
    ctx->host_graph = false;
    ctx->cache_bytes = 0;
}

int gw_stream_prepare(gw_ctx* ctx) {
    if (!ctx->slot_mem.empty()) return GW_OK;
    const uint32_t nb = ctx->nblocks;
    // a slot holds the largest block: its offsets, then (256-byte aligned)
    // its columns
    uint64_t slot = 0;
    for (uint32_t b = 0; b < nb; ++b) slot = std::max(slot, align_up(block_bytes(ctx, b, nullptr), kSlotAlign));
    const uint64_t nslots = std::min<uint64_t>(nb, ctx->cache_bytes / slot);
    if (nslots == 0)
        return gw_fail(ctx, GW_ENOMEM, "gw_run: the block cache (" + std::to_string(ctx->cache_bytes) +
                                           " bytes) is smaller than one block slot (" + std::to_string(slot) + " bytes)");
    ctx->slot_bytes = slot;
    for (uint64_t i = 0; i < nslots; ++i) {
        void* m = nullptr;
        int rc = gw_alloc(ctx, &m, slot, "block cache slot");
        if (rc) {
            gw_stream_reset(ctx);
            return rc;
        }
        ctx->slot_mem.push_back(m);
    }
    ctx->slot_block.assign(nslots, -1);
    ctx->block_slot.assign(nb, -1);
    GW_HIP(ctx, hipHostMalloc(&ctx->h_views, nb * 16ull, hipHostMallocDefault));
    std::memset(ctx->h_views, 0, nb * 16ull);
    int rc = gw_alloc(ctx, &ctx->d_views, nb * 16ull, "block views");
    if (rc) return rc;
    if (!ctx->copy_stream) GW_HIP(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    for (auto& e : ctx->copy_ev)
        if (!e) GW_HIP(ctx, hipEventCreate(&e));
    return GW_OK;
}

// The CSR in ctx->h_off / ctx->h_col (pinned) becomes the context's graph:
// vertex counts and the default block table of an out-of-core run.
int gw_host_graph_adopt(gw_ctx* ctx, uint32_t n, uint64_t m, uint64_t cache_bytes) {
    ctx->host_graph = true;
    ctx->n = n;
    ctx->m = m;
    ctx->cache_bytes = cache_bytes;
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_counts, n * 4ull, "vertex counts"))) {
        gw_host_graph_free(ctx);
        return rc;
    }
    // default block table: blocks of at most cache_bytes / 4 (at least four
    // slots), grown until it has at most 2048 blocks
    uint64_t target = std::max<uint64_t>(cache_bytes / 4, 64);
    std::vector<uint32_t> bf = cut_blocks(ctx, target);
    while (bf.size() - 1 > 2048) {
        target *= 2;
        bf = cut_blocks(ctx, target);
    }
    return gw_graph_blocks(ctx, bf.data(), static_cast<uint32_t>(bf.size() - 1));
}

extern "C" {

int gw_host_alloc(gw_ctx* ctx, uint64_t bytes, void** out) {
    if (!ctx || !out || bytes == 0) return gw_fail(ctx, GW_EINVAL, "gw_host_alloc: bad arguments");
    *out = nullptr;
    GW_HIP(ctx, hipSetDevice(ctx->device));
    hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // a refused allocation must not surface at the next launch check
        *out = nullptr;
        return gw_fail(ctx, GW_ENOMEM, "gw_host_alloc: hipHostMalloc(" + std::to_string(bytes) + ") failed: " +
                                           hipGetErrorString(e));
    }
    ctx->pinned.emplace_back(*out, bytes);
    return GW_OK;
}

int gw_host_free(gw_ctx* ctx, void* p) {
    if (!ctx || !p) return gw_fail(ctx, GW_EINVAL, "gw_host_free: bad arguments");
    auto it = std::find_if(ctx->pinned.begin(), ctx->pinned.end(), [&](const auto& a) { return a.first == p; });
    if (it == ctx->pinned.end()) return gw_fail(ctx, GW_EINVAL, "gw_host_free: not a gw_host_alloc pointer");
    const char* b = static_cast<const char*>(p);
    auto inside = [&](const void* q) {
        const char* c = static_cast<const char*>(q);
        return q && c >= b && c < b + it->second;
    };
    if (ctx->host_graph && (inside(ctx->h_off) || inside(ctx->h_col)))
        return gw_fail(ctx, GW_ESTATE, "gw_host_free: the host-resident graph still uses this memory");
    GW_HIP(ctx, hipHostFree(p));
    ctx->pinned.erase(it);
    return GW_OK;
}

int gw_graph_set_host(gw_ctx* ctx, const uint64_t* off, const uint32_t* col, uint32_t n, uint64_t m,
                      uint64_t cache_bytes) {
    if (!ctx || !off || (m && !col) || n == 0 || n == 0xFFFFFFFFu || cache_bytes == 0)
        return gw_fail(ctx, GW_EINVAL, "gw_graph_set_host: bad arguments");
    if (off[0] != 0 || off[n] != m) return gw_fail(ctx, GW_EINVAL, "gw_graph_set_host: off[0] must be 0 and off[n] == m");
    for (uint32_t v = 0; v < n; ++v)
        if (off[v + 1] < off[v]) return gw_fail(ctx, GW_EINVAL, "gw_graph_set_host: offsets not monotone");
    for (uint64_t e = 0; e < m; ++e)
        if (col[e] >= n) return gw_fail(ctx, GW_EINVAL, "gw_graph_set_host: column out of range");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_graph_free(ctx);
    // pinned memory for the copies: the caller's arrays are used in place
    // when gw_host_alloc made them
    const uint64_t ob = (n + 1ull) * 8, cb = m * 4;
    if (gw_is_pinned(ctx, off, ob)) {
        ctx->h_off = off;
    } else {
        GW_HIP(ctx, hipHostMalloc(&ctx->h_owned_off, ob, hipHostMallocDefault));
        std::memcpy(ctx->h_owned_off, off, ob);
        ctx->h_off = static_cast<const uint64_t*>(ctx->h_owned_off);
    }
    if (m == 0 || gw_is_pinned(ctx, col, cb)) {
        ctx->h_col = col;
    } else {
        hipError_t e = hipHostMalloc(&ctx->h_owned_col, cb, hipHostMallocDefault);
        if (e != hipSuccess) {
            gw_host_graph_free(ctx);
            return gw_hip_check(ctx, e, "hipHostMalloc(columns)");
        }
        std::memcpy(ctx->h_owned_col, col, cb);
        ctx->h_col = static_cast<const uint32_t*>(ctx->h_owned_col);
    }
    return gw_host_graph_adopt(ctx, n, m, cache_bytes);
}

// Random-read ceiling of pinned host memory from a kernel: `gathers`
// independent 4-byte reads over a pinned buffer of `bytes` bytes (kernel time
// in ms, after a warm-up of an eighth as many).
int gw_host_gather_probe(gw_ctx* ctx, uint64_t bytes, uint64_t gathers, double* ms) {
    if (!ctx || bytes < 4096 || gathers == 0 || bytes / 4 > 0xFFFFFFFFull)
        return gw_fail(ctx, GW_EINVAL, "gw_host_gather_probe: bytes in [4096, 16 GiB], gathers > 0");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    void* h = nullptr;
    GW_HIP(ctx, hipHostMalloc(&h, bytes, hipHostMallocDefault));
    std::memset(h, 1, bytes);
    uint32_t* out = nullptr;
    const uint32_t tb = 256, grid = std::max(1u, ctx->cu_count) * 64;
    int rc = gw_alloc(ctx, (void**)&out, static_cast<uint64_t>(grid) * tb * 4, "host gather out");
    if (rc) {
        (void)hipHostFree(h);
        return rc;
    }
    const uint32_t nwords = static_cast<uint32_t>(bytes / 4);
    host_gather_kernel<<<dim3(grid), dim3(tb), 0, ctx->stream>>>(static_cast<const uint32_t*>(h), nwords, gathers / 8 + 1, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[0], ctx->stream);
    if (e == hipSuccess) {
        host_gather_kernel<<<dim3(grid), dim3(tb), 0, ctx->stream>>>(static_cast<const uint32_t*>(h), nwords, gathers, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[1]);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]);
    gw_free(ctx, out, static_cast<uint64_t>(grid) * tb * 4);
    (void)hipHostFree(h);
    if (e != hipSuccess) return gw_hip_check(ctx, e, "host gather probe");
    if (ms) *ms = t;
    return GW_OK;
}

// Host-to-device copy ceiling: `bytes` from pinned host memory into HBM on
// the copy stream (the path block loads take); copy time in ms.
int gw_h2d_probe(gw_ctx* ctx, uint64_t bytes, double* ms) {
    if (!ctx || bytes == 0) return gw_fail(ctx, GW_EINVAL, "gw_h2d_probe: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->copy_stream) GW_HIP(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    for (auto& e : ctx->copy_ev)
        if (!e) GW_HIP(ctx, hipEventCreate(&e));
    void* h = nullptr;
    void* d = nullptr;
    GW_HIP(ctx, hipHostMalloc(&h, bytes, hipHostMallocDefault));
    std::memset(h, 1, bytes);
    int rc = gw_alloc(ctx, &d, bytes, "h2d probe");
    if (rc) {
        (void)hipHostFree(h);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->copy_stream);  // warm
    if (e == hipSuccess) e = hipEventRecord(ctx->copy_ev[1], ctx->copy_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->copy_ev[2], ctx->copy_stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->copy_ev[2]);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->copy_ev[1], ctx->copy_ev[2]);
    gw_free(ctx, d, bytes);
    (void)hipHostFree(h);
    if (e != hipSuccess) return gw_hip_check(ctx, e, "h2d probe");
    if (ms) *ms = t;
    return GW_OK;
}

}  // extern "C"
