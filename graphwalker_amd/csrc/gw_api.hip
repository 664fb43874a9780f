// gw_api.hip -- the C ABI (include/gwhip.h): context, graph and walk residency,
// run dispatch, result read-back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "gw_device.h"
#include "gw_internal.h"

int gw_fail(gw_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int gw_hip_check(gw_ctx* ctx, hipError_t e, const char* what) {
    return gw_fail(ctx, e == hipErrorOutOfMemory ? GW_ENOMEM : GW_EHIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

int gw_alloc(gw_ctx* ctx, void** p, uint64_t bytes, const char* what) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    if (ctx->hbm_budget && ctx->hbm_used + bytes > ctx->hbm_budget)
        return gw_fail(ctx, GW_ENOMEM, std::string(what) + ": HBM budget exceeded (" +
                                           std::to_string(ctx->hbm_used + bytes) + " > " +
                                           std::to_string(ctx->hbm_budget) + " bytes)");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return gw_fail(ctx, GW_ENOMEM, std::string(what) + ": hipMalloc(" + std::to_string(bytes) +
                                           ") failed: " + hipGetErrorString(e));
    }
    ctx->hbm_used += bytes;
    return GW_OK;
}

void gw_free(gw_ctx* ctx, void* p, uint64_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    if (bytes == 0) bytes = 16;
    ctx->hbm_used = ctx->hbm_used >= bytes ? ctx->hbm_used - bytes : 0;
}

void gw_graph_free(gw_ctx* ctx) {
    ++ctx->graph_epoch;
    gw_host_graph_free(ctx);
    gw_free(ctx, ctx->d_off, (ctx->n + 1ull) * 8);
    gw_free(ctx, ctx->d_col, ctx->m * 4);
    gw_free(ctx, ctx->d_counts, ctx->n * 4ull);
    gw_free(ctx, ctx->d_block_first, (ctx->nblocks + 1ull) * 4);
    gw_free(ctx, ctx->d_ext, std::max<uint64_t>(1, ctx->m) * 16);
    ctx->d_ext = nullptr;
    ctx->ext_failed = false;
    ctx->d_off = nullptr;
    ctx->d_col = nullptr;
    ctx->d_counts = nullptr;
    ctx->d_block_first = nullptr;
    ctx->n = 0;
    ctx->m = 0;
    ctx->nblocks = 0;
    ctx->block_first.clear();
}

namespace {

}  // namespace

// The walk array no longer holds the walks of gw_walks_init_sources (the
// source arrays stay allocated for the next placement).
void gw_gen_clear(gw_ctx* ctx) {
    ctx->gen_nsrc = ctx->gen_wps = 0;
    ctx->gen_built[0] = ~0ull;
}

void gw_gen_free(gw_ctx* ctx) {
    gw_gen_clear(ctx);
    gw_free(ctx, ctx->d_gen_src, ctx->gen_cap * 4ull);
    gw_free(ctx, ctx->d_gen_field, ctx->gen_cap * 4ull);
    gw_free(ctx, ctx->d_gen_table, ctx->gen_cap * sizeof(gw::SourceStart));
    ctx->d_gen_src = ctx->d_gen_field = nullptr;
    ctx->d_gen_table = nullptr;
    ctx->gen_cap = 0;
}

namespace {

void free_walks(gw_ctx* ctx) {
    gw_gen_free(ctx);
    gw_free(ctx, ctx->d_walks, ctx->walk_cap * sizeof(gw_walk));
    gw_free(ctx, ctx->d_walks_alt, ctx->walk_cap * sizeof(gw_walk));
    gw_free(ctx, ctx->d_final, ctx->walk_cap * sizeof(gw_walk));
    gw_free(ctx, ctx->d_work, ctx->walk_cap * sizeof(gw_walk));
    ctx->d_work = nullptr;
    gw_free(ctx, ctx->d_hash, ctx->walk_cap * 8);
    gw_free(ctx, ctx->d_visits, ctx->walk_cap * 4);
    gw_free(ctx, ctx->d_events, ctx->walk_cap * ctx->events_cap * 8);
    ctx->d_events = nullptr;
    ctx->events_cap = 0;
    ctx->d_walks = ctx->d_walks_alt = ctx->d_final = nullptr;
    ctx->d_hash = nullptr;
    ctx->d_visits = nullptr;
    ctx->walk_cap = 0;
    ctx->nwalks = 0;
}

int reserve_walks(gw_ctx* ctx, uint64_t n) {
    gw_pool_free(ctx);  // device walks replace a host pool
    gw_gen_clear(ctx);
    if (n <= ctx->walk_cap) {
        ctx->nwalks = n;
        return GW_OK;
    }
    free_walks(ctx);
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_walks, n * sizeof(gw_walk), "walks"))) return rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_walks_alt, n * sizeof(gw_walk), "walks (alt)")) ||
        (rc = gw_alloc(ctx, (void**)&ctx->d_final, n * sizeof(gw_walk), "final walks"))) {
        gw_free(ctx, ctx->d_walks, n * sizeof(gw_walk));
        gw_free(ctx, ctx->d_walks_alt, n * sizeof(gw_walk));
        ctx->d_walks = ctx->d_walks_alt = nullptr;
        return rc;
    }
    ctx->walk_cap = n;
    ctx->nwalks = n;
    return GW_OK;
}

int single_block(gw_ctx* ctx) {
    ctx->nblocks = 1;
    ctx->block_first = {0u, ctx->n};
    int rc = gw_alloc(ctx, (void**)&ctx->d_block_first, 2 * 4, "block table");
    if (rc) return rc;
    GW_HIP(ctx, hipMemcpy(ctx->d_block_first, ctx->block_first.data(), 8, hipMemcpyHostToDevice));
    return GW_OK;
}

__global__ void init_per_vertex_kernel(gw_walk* __restrict__ w, uint64_t nw, uint32_t per_vertex) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= nw) return;
    const uint32_t v = static_cast<uint32_t>(i / per_vertex);
    w[i] = gw_walk{static_cast<uint32_t>(i), v, v, 0u};
}

__global__ void init_sources_kernel(gw_walk* __restrict__ w, uint64_t nw, const uint32_t* __restrict__ src,
                                    const uint32_t* __restrict__ field, uint32_t wps) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= nw) return;
    const uint32_t k = static_cast<uint32_t>(i / wps);
    const uint32_t v = src[k];
    w[i] = gw_walk{static_cast<uint32_t>(i), field ? field[k] : v, v, 0u};
}

// Independent random 4-byte gathers: index i's address comes from Philox(i),
// not from loaded data, so the loads of a lane are all in flight together.
__global__ void gather_probe_kernel(const uint32_t* __restrict__ buf, uint32_t nwords, uint64_t gathers,
                                    uint32_t* __restrict__ out) {
    const uint64_t tid = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    uint32_t acc = 0;
    for (uint64_t i = tid; i < gathers; i += stride) {
        const gw::u32x4 r = gw::philox4x32_10(static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32),
                                              0x47415448u, 0, 1, 2);
        acc ^= buf[gw::mulhi32(r.x, nwords)];
    }
    out[tid] = acc;
}

// Streaming copy of `n16` 16-byte words: the known-size kernel that hardware
// byte counters are calibrated against.
__global__ void copy_probe_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, uint64_t n16) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n16; i += stride) b[i] = a[i];
}

}  // namespace

int gw_reserve_walks(gw_ctx* ctx, uint64_t n) { return reserve_walks(ctx, n); }

extern "C" {

int gw_copy_probe(gw_ctx* ctx, uint64_t bytes, double* ms) {
    if (!ctx || bytes < 4096 || bytes % 16) return gw_fail(ctx, GW_EINVAL, "gw_copy_probe: bytes must be >= 4096 and a multiple of 16");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    uint4 *a = nullptr, *b = nullptr;
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&a, bytes, "copy src"))) return rc;
    if ((rc = gw_alloc(ctx, (void**)&b, bytes, "copy dst"))) { gw_free(ctx, a, bytes); return rc; }
    const uint32_t tb = 256, grid = std::max(1u, ctx->cu_count) * 16;
    hipError_t e = hipMemsetAsync(a, 3, bytes, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[0], ctx->stream);
    if (e == hipSuccess) {
        copy_probe_kernel<<<dim3(grid), dim3(tb), 0, ctx->stream>>>(a, b, bytes / 16);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[1]);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]);
    gw_free(ctx, a, bytes);
    gw_free(ctx, b, bytes);
    if (e != hipSuccess) return gw_hip_check(ctx, e, "copy probe");
    if (ms) *ms = t;
    return GW_OK;
}

int gw_create(gw_ctx** out, const gw_cfg* cfg) {
    if (!out) return GW_EINVAL;
    *out = nullptr;
    gw_ctx* ctx = new gw_ctx();
    ctx->device = cfg ? cfg->device : 0;
    ctx->hbm_budget = cfg ? cfg->hbm_budget_bytes : 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= ctx->device || ctx->device < 0) {
        delete ctx;
        return GW_EHIP;
    }
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return GW_EHIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
    for (auto& e : ctx->ev) (void)hipEventCreate(&e);
    if (gw_alloc(ctx, (void**)&ctx->d_counters, kCounters * sizeof(unsigned long long), "counters")) {
        gw_destroy(ctx);
        return GW_ENOMEM;
    }
    // the device address of gw_run_noalloc's static scratch, with this device current
    if (gw_noalloc_init(ctx) != GW_OK) {
        gw_destroy(ctx);
        return GW_EHIP;
    }
    *out = ctx;
    return GW_OK;
}

void gw_destroy(gw_ctx* ctx) {
    if (!ctx) return;
    gw_serve_close(ctx);
    gw_run_graph_free(ctx);
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    gw_graph_free(ctx);
    gw_pool_free(ctx);
    free_walks(ctx);
    gw_free(ctx, ctx->d_counters, kCounters * sizeof(unsigned long long));
    if (ctx->h_run_cnt) (void)hipHostFree(ctx->h_run_cnt);
    gw_free(ctx, ctx->d_tile_hist, ctx->tile_hist_cap * 4);
    gw_free(ctx, ctx->d_table, ctx->table_bytes);
    gw_free(ctx, ctx->d_keys, ctx->keys_cap * 8);
    gw_free(ctx, ctx->d_keys_alt, ctx->runs_cap * 8);
    gw_free(ctx, ctx->d_sort_temp, ctx->sort_temp_bytes);
    gw_free(ctx, ctx->d_part_keys, ctx->part_keys_cap * 4);
    gw_free(ctx, ctx->d_part_counts, ctx->part_keys_cap * 4);
    gw_free(ctx, ctx->d_part_off, (ctx->part_fields_cap + 1ull) * 8);
    gw_free(ctx, ctx->d_part_cursor, ctx->part_fields_cap * 8ull);
    gw_free(ctx, ctx->d_part_pairs, ctx->part_fields_cap * 8ull);
    gw_free(ctx, ctx->d_part_fvert, ctx->part_fields_cap * 4ull);
    gw_free(ctx, ctx->d_part_fcount, ctx->part_fields_cap * 4ull);
    gw_free(ctx, ctx->d_part_items, ctx->part_fields_cap * 32ull);
    gw_free(ctx, ctx->d_split_counts, ctx->split_cap * 8);
    gw_free(ctx, ctx->d_runs, ctx->runs_cap * 8);
    gw_free(ctx, ctx->d_run_counts, ctx->runs_cap * 4);
    gw_free(ctx, ctx->d_bucket_off, 0);
    gw_free(ctx, ctx->d_block_minhop, ctx->block_minhop_cap * 4ull);
    gw_free(ctx, ctx->d_meet, ctx->meet_bytes);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->copy_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& a : ctx->pinned) (void)hipHostFree(a.first);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* gw_last_error(gw_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int gw_device_info(gw_ctx* ctx, char* name, size_t name_len, uint32_t* cu_count, uint64_t* hbm_bytes) {
    if (!ctx) return GW_EINVAL;
    hipDeviceProp_t prop;
    GW_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return GW_OK;
}

int gw_graph_set(gw_ctx* ctx, const uint64_t* off, const uint32_t* col, uint32_t n, uint64_t m) {
    if (!ctx || !off || (m && !col) || n == 0 || n == 0xFFFFFFFFu) return gw_fail(ctx, GW_EINVAL, "gw_graph_set: bad arguments");
    if (off[0] != 0 || off[n] != m) return gw_fail(ctx, GW_EINVAL, "gw_graph_set: off[0] must be 0 and off[n] == m");
    for (uint32_t v = 0; v < n; ++v)
        if (off[v + 1] < off[v]) return gw_fail(ctx, GW_EINVAL, "gw_graph_set: offsets not monotone");
    for (uint64_t e = 0; e < m; ++e)
        if (col[e] >= n) return gw_fail(ctx, GW_EINVAL, "gw_graph_set: column out of range");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_graph_free(ctx);
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_off, (n + 1ull) * 8, "offsets"))) return rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_col, m * 4, "columns"))) return rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_counts, n * 4ull, "vertex counts"))) return rc;
    ctx->n = n;
    ctx->m = m;
    GW_HIP(ctx, hipMemcpy(ctx->d_off, off, (n + 1ull) * 8, hipMemcpyHostToDevice));
    if (m) GW_HIP(ctx, hipMemcpy(ctx->d_col, col, m * 4, hipMemcpyHostToDevice));
    return single_block(ctx);
}

int gw_graph_set_edges(gw_ctx* ctx, const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t n,
                       uint32_t flags) {
    if (!ctx || (m && (!src || !dst)) || (flags & ~(GW_EDGES_DROP_SELF_LOOPS | GW_EDGES_SORT_ROWS)) ||
        n == 0xFFFFFFFFu)
        return gw_fail(ctx, GW_EINVAL, "gw_graph_set_edges: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_graph_free(ctx);
    int rc = gw_edges_build(ctx, src, dst, m, n, flags);
    if (rc) return rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_counts, ctx->n * 4ull, "vertex counts"))) return rc;
    return single_block(ctx);
}

int gw_graph_rmat(gw_ctx* ctx, uint32_t scale, uint64_t m, uint64_t seed, uint32_t ta, uint32_t tab,
                  uint32_t tabc) {
    if (!ctx || scale == 0 || scale > 31 || !(ta <= tab && tab <= tabc && tabc <= 65536))
        return gw_fail(ctx, GW_EINVAL, "gw_graph_rmat: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_graph_free(ctx);
    int rc = gw_rmat_build(ctx, scale, m, seed, ta, tab, tabc);
    if (rc) return rc;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_counts, ctx->n * 4ull, "vertex counts"))) return rc;
    return single_block(ctx);
}

int gw_graph_rmat_host(gw_ctx* ctx, uint32_t scale, uint64_t m, uint64_t seed, uint32_t ta, uint32_t tab,
                       uint32_t tabc, uint64_t cache_bytes, uint64_t work_bytes, double* build_ms) {
    if (!ctx || scale == 0 || scale > 31 || cache_bytes == 0 || !(ta <= tab && tab <= tabc && tabc <= 65536))
        return gw_fail(ctx, GW_EINVAL, "gw_graph_rmat_host: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_graph_free(ctx);
    return gw_rmat_build_host(ctx, scale, m, seed, ta, tab, tabc, cache_bytes, work_bytes, build_ms);
}

int gw_graph_transpose(gw_ctx* ctx) {
    if (!ctx || !ctx->d_off) return gw_fail(ctx, GW_ESTATE, "gw_graph_transpose: needs a device-resident graph");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ++ctx->graph_epoch;
    // the successor-header edges describe the old edges
    gw_free(ctx, ctx->d_ext, std::max<uint64_t>(1, ctx->m) * 16);
    ctx->d_ext = nullptr;
    ctx->ext_failed = false;
    return gw_transpose_build(ctx);
}

int gw_graph_blocks(gw_ctx* ctx, const uint32_t* block_first, uint32_t nblocks) {
    if (!ctx || !gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_graph_blocks: no graph");
    if (!block_first || nblocks == 0) return gw_fail(ctx, GW_EINVAL, "gw_graph_blocks: empty block table");
    if (nblocks > 2048) return gw_fail(ctx, GW_EINVAL, "gw_graph_blocks: at most 2048 blocks");
    if (block_first[0] != 0 || block_first[nblocks] != ctx->n)
        return gw_fail(ctx, GW_EINVAL, "gw_graph_blocks: table must start at 0 and end at n");
    for (uint32_t b = 0; b < nblocks; ++b)
        if (block_first[b + 1] <= block_first[b]) return gw_fail(ctx, GW_EINVAL, "gw_graph_blocks: empty or unordered block");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    gw_stream_reset(ctx);  // the block cache is sized for the old table
    gw_free(ctx, ctx->d_block_first, (ctx->nblocks + 1ull) * 4);
    ctx->d_block_first = nullptr;
    int rc = gw_alloc(ctx, (void**)&ctx->d_block_first, (nblocks + 1ull) * 4, "block table");
    if (rc) return rc;
    ctx->nblocks = nblocks;
    ctx->block_first.assign(block_first, block_first + nblocks + 1);
    GW_HIP(ctx, hipMemcpy(ctx->d_block_first, block_first, (nblocks + 1ull) * 4, hipMemcpyHostToDevice));
    return GW_OK;
}

int gw_graph_info(gw_ctx* ctx, uint32_t* n, uint64_t* m, uint32_t* nblocks) {
    if (!ctx) return GW_EINVAL;
    if (n) *n = ctx->n;
    if (m) *m = ctx->m;
    if (nblocks) *nblocks = ctx->nblocks;
    return GW_OK;
}

int gw_graph_read(gw_ctx* ctx, uint64_t* off, uint32_t* col) {
    if (!ctx || !gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_graph_read: no graph");
    if (ctx->host_graph) {
        if (off) std::memcpy(off, ctx->h_off, (ctx->n + 1ull) * 8);
        if (col && ctx->m) std::memcpy(col, ctx->h_col, ctx->m * 4);
        return GW_OK;
    }
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (off) GW_HIP(ctx, hipMemcpy(off, ctx->d_off, (ctx->n + 1ull) * 8, hipMemcpyDeviceToHost));
    if (col && ctx->m) GW_HIP(ctx, hipMemcpy(col, ctx->d_col, ctx->m * 4, hipMemcpyDeviceToHost));
    return GW_OK;
}

int gw_walks_load(gw_ctx* ctx, const gw_walk* walks, uint64_t n) {
    if (!ctx || (n && !walks)) return gw_fail(ctx, GW_EINVAL, "gw_walks_load: bad arguments");
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_walks_load: no graph");
    if (n >= 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_walks_load: walk ids are 32-bit");
    uint32_t src_max = 0, v_max = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (walks[i].id >= n || walks[i].vertex >= ctx->n)
            return gw_fail(ctx, GW_EINVAL, "gw_walks_load: walk " + std::to_string(i) + " has id >= n or vertex out of range");
        src_max = std::max(src_max, walks[i].source);
        v_max = std::max(v_max, walks[i].vertex);
    }
    GW_HIP(ctx, hipSetDevice(ctx->device));
    int rc = reserve_walks(ctx, n);
    if (rc) return rc;
    if (n) GW_HIP(ctx, hipMemcpy(ctx->d_walks, walks, n * sizeof(gw_walk), hipMemcpyHostToDevice));
    ctx->walk_src_max = src_max;
    ctx->walk_v_max = v_max;
    ctx->field_walks.clear();
    ctx->field_vertex.clear();
    ctx->part_per_walk = 0;
    if (n && src_max < kPartFieldLimit) {
        ctx->field_walks.assign(src_max + 1ull, 0);
        ctx->field_vertex.assign(src_max + 1ull, 0xFFFFFFFFu);
        for (uint64_t i = 0; i < n; ++i) {
            if (ctx->field_walks[walks[i].source]++ == 0) ctx->field_vertex[walks[i].source] = walks[i].vertex;
        }
    }
    return GW_OK;
}

int gw_walks_init_per_vertex(gw_ctx* ctx, uint32_t per_vertex) {
    if (!ctx || per_vertex == 0) return gw_fail(ctx, GW_EINVAL, "gw_walks_init_per_vertex: bad arguments");
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_walks_init_per_vertex: no graph");
    const uint64_t nw = static_cast<uint64_t>(ctx->n) * per_vertex;
    if (nw >= 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_walks_init_per_vertex: walk ids are 32-bit");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    int rc = reserve_walks(ctx, nw);
    if (rc) return rc;
    const uint32_t tb = 256;
    init_per_vertex_kernel<<<dim3(static_cast<uint32_t>((nw + tb - 1) / tb)), dim3(tb), 0, ctx->stream>>>(ctx->d_walks, nw, per_vertex);
    GW_HIP(ctx, hipGetLastError());
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->walk_src_max = ctx->n - 1;
    ctx->walk_v_max = ctx->n - 1;
    ctx->field_walks.clear();
    ctx->field_vertex.clear();
    ctx->part_per_walk = 0;
    if (ctx->n <= kPartFieldLimit) {
        ctx->field_walks.assign(ctx->n, per_vertex);
        ctx->field_vertex.resize(ctx->n);
        for (uint32_t v = 0; v < ctx->n; ++v) ctx->field_vertex[v] = v;  // walk i starts at vertex i / per_vertex
    }
    return GW_OK;
}

int gw_walks_init_sources(gw_ctx* ctx, const uint32_t* sources, const uint32_t* source_field, uint32_t nsources,
                          uint32_t walks_per_source) {
    if (!ctx || !sources || nsources == 0 || walks_per_source == 0)
        return gw_fail(ctx, GW_EINVAL, "gw_walks_init_sources: bad arguments");
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_walks_init_sources: no graph");
    uint32_t src_max = 0, v_max = 0;
    for (uint32_t k = 0; k < nsources; ++k) {
        if (sources[k] >= ctx->n) return gw_fail(ctx, GW_EINVAL, "gw_walks_init_sources: source out of range");
        src_max = std::max(src_max, source_field ? source_field[k] : sources[k]);
        v_max = std::max(v_max, sources[k]);
    }
    const uint64_t nw = static_cast<uint64_t>(nsources) * walks_per_source;
    if (nw >= 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_walks_init_sources: walk ids are 32-bit");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    int rc = reserve_walks(ctx, nw);
    if (rc) return rc;
    // the sources stay on the device: walk starts are generated from them
    // (the walk records are written too, for every other path)
    // (kept allocated across placements, so a captured run can replay new sources)
    if (nsources > ctx->gen_cap) {
        gw_gen_free(ctx);
        if ((rc = gw_alloc(ctx, (void**)&ctx->d_gen_src, nsources * 4ull, "sources")) ||
            (rc = gw_alloc(ctx, (void**)&ctx->d_gen_field, nsources * 4ull, "source fields")) ||
            (rc = gw_alloc(ctx, &ctx->d_gen_table, nsources * sizeof(gw::SourceStart), "source start table"))) {
            gw_gen_free(ctx);
            return rc;
        }
        ctx->gen_cap = nsources;
    }
    ctx->gen_nsrc = nsources;
    GW_HIP(ctx, hipMemcpy(ctx->d_gen_src, sources, nsources * 4ull, hipMemcpyHostToDevice));
    GW_HIP(ctx, hipMemcpy(ctx->d_gen_field, source_field ? source_field : sources, nsources * 4ull,
                          hipMemcpyHostToDevice));
    const uint32_t tb = 256;
    init_sources_kernel<<<dim3(static_cast<uint32_t>((nw + tb - 1) / tb)), dim3(tb), 0, ctx->stream>>>(
        ctx->d_walks, nw, ctx->d_gen_src, ctx->d_gen_field, walks_per_source);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        gw_gen_clear(ctx);
        return gw_hip_check(ctx, e, "init_sources_kernel");
    }
    ctx->gen_wps = walks_per_source;
    ctx->walk_src_max = src_max;
    ctx->walk_v_max = v_max;
    ctx->field_walks.clear();
    ctx->field_vertex.clear();
    ctx->part_per_walk = 0;
    if (src_max < kPartFieldLimit) {
        ctx->field_walks.assign(src_max + 1ull, 0);
        ctx->field_vertex.assign(src_max + 1ull, 0xFFFFFFFFu);
        for (uint32_t k = 0; k < nsources; ++k) {
            const uint32_t f = source_field ? source_field[k] : sources[k];
            if (ctx->field_walks[f] == 0) ctx->field_vertex[f] = sources[k];
            ctx->field_walks[f] += walks_per_source;
        }
    }
    return GW_OK;
}

int gw_walks_read(gw_ctx* ctx, gw_walk* out, uint64_t n) {
    if (!ctx || (n && !out)) return gw_fail(ctx, GW_EINVAL, "gw_walks_read: bad arguments");
    if (ctx->pool) {
        const gw_pool& P = *ctx->pool;
        if (!P.keep_final || !P.final_valid || n > P.n)
            return gw_fail(ctx, GW_ESTATE, "gw_walks_read: the host pool keeps no final records (GW_POOL_KEEP_FINAL) or no run yet");
        std::copy(P.h_final, P.h_final + n, out);
        return GW_OK;
    }
    if (n > ctx->nwalks) return gw_fail(ctx, GW_EINVAL, "gw_walks_read: more records requested than loaded");
    if (!ctx->finals_valid) return gw_fail(ctx, GW_ESTATE, "gw_walks_read: the last run kept no final records (GW_REC_NO_FINAL)");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (n) GW_HIP(ctx, hipMemcpy(out, ctx->d_final, n * sizeof(gw_walk), hipMemcpyDeviceToHost));
    return GW_OK;
}

int gw_walks_current(gw_ctx* ctx, gw_walk* out, uint64_t n) {
    if (!ctx || (n && !out)) return gw_fail(ctx, GW_EINVAL, "gw_walks_current: bad arguments");
    if (ctx->pool) return gw_fail(ctx, GW_ESTATE, "gw_walks_current: the walks are in a host pool");
    if (n > ctx->nwalks) return gw_fail(ctx, GW_EINVAL, "gw_walks_current: more records requested than loaded");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (n) GW_HIP(ctx, hipMemcpy(out, ctx->d_walks, n * sizeof(gw_walk), hipMemcpyDeviceToHost));
    return GW_OK;
}

int gw_run(gw_ctx* ctx, const gw_program* prog, gw_stats* stats) {
    if (!ctx || !prog) return gw_fail(ctx, GW_EINVAL, "gw_run: bad arguments");
    if (!gw_has_graph(ctx)) return gw_fail(ctx, GW_ESTATE, "gw_run: no graph");
    if (prog->kind > GW_RESTART) return gw_fail(ctx, GW_EINVAL, "gw_run: unknown program kind");
    if (prog->L == 0 && prog->kind != GW_PROB) return gw_fail(ctx, GW_EINVAL, "gw_run: L must be >= 1");
    if (prog->stop_thr > (1ull << 32)) return gw_fail(ctx, GW_EINVAL, "gw_run: stop_thr above 2^32");
    if (prog->kind == GW_JUMP && prog->nvertices == 0) return gw_fail(ctx, GW_EINVAL, "gw_run: JUMP needs nvertices");
    if (prog->kind == GW_JUMP && prog->nvertices > ctx->n) return gw_fail(ctx, GW_EINVAL, "gw_run: nvertices above graph size");
    if (prog->recorder & ~(GW_REC_VERTEX_COUNT | GW_REC_WALK_SUMMARY | GW_REC_EVENTS | GW_REC_SOURCE_COUNT | GW_REC_NO_FINAL |
                           GW_REC_EVENTS_LAST))
        return gw_fail(ctx, GW_EINVAL, "gw_run: unknown recorder");
    if ((prog->recorder & GW_REC_EVENTS_LAST) && !(prog->recorder & GW_REC_EVENTS))
        return gw_fail(ctx, GW_EINVAL, "gw_run: GW_REC_EVENTS_LAST needs GW_REC_EVENTS");
    if ((prog->recorder & GW_REC_SOURCE_COUNT) && (prog->recorder & (GW_REC_VERTEX_COUNT | GW_REC_EVENTS)))
        return gw_fail(ctx, GW_EINVAL, "gw_run: GW_REC_SOURCE_COUNT combines with GW_REC_WALK_SUMMARY only");
    if (prog->source_method > GW_SRC_PARTITIONED) return gw_fail(ctx, GW_EINVAL, "gw_run: unknown source_method");
    if ((prog->recorder & GW_REC_EVENTS) && prog->event_cap == 0)
        return gw_fail(ctx, GW_EINVAL, "gw_run: GW_REC_EVENTS needs event_cap > 0");
    // Restart programs jump to the source field, which must then name a vertex.
    if (ctx->nwalks && (prog->kind == GW_RESTART || prog->kind == GW_RESTART_JOINT) && ctx->walk_src_max >= ctx->n)
        return gw_fail(ctx, GW_EINVAL, "gw_run: restart program with a source field that is not a vertex");
    // An uncapped program must end: require a stop probability of at least 2^-12.
    if (prog->kind == GW_PROB && prog->stop_thr < (1ull << 20))
        return gw_fail(ctx, GW_EINVAL, "gw_run: PROB needs stop_thr >= 2^20 (mean walk length <= 4096)");
    if (ctx->nwalks && ctx->walk_v_max >= ctx->n)
        return gw_fail(ctx, GW_ESTATE, "gw_run: walks start at vertices outside the graph (load walks again after changing the graph)");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    return gw_step_run(ctx, prog, stats);
}

int gw_counts_read(gw_ctx* ctx, uint32_t* counts, uint32_t n) {
    if (!ctx || !counts) return gw_fail(ctx, GW_EINVAL, "gw_counts_read: bad arguments");
    if (!ctx->d_counts || n != ctx->n) return gw_fail(ctx, GW_EINVAL, "gw_counts_read: n must equal the graph's vertex count");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    GW_HIP(ctx, hipMemcpy(counts, ctx->d_counts, n * 4ull, hipMemcpyDeviceToHost));
    return GW_OK;
}

int gw_summary_read(gw_ctx* ctx, uint64_t* hash, uint32_t* visits, uint64_t n) {
    if (!ctx) return GW_EINVAL;
    if (ctx->pool) {
        const gw_pool& P = *ctx->pool;
        if (!P.summary_valid || n > P.n) return gw_fail(ctx, GW_ESTATE, "gw_summary_read: no summary recorded for that many walks");
        if (hash) std::copy(P.h_hash, P.h_hash + n, hash);
        if (visits) std::copy(P.h_visits, P.h_visits + n, visits);
        return GW_OK;
    }
    if (!ctx->d_hash || n > ctx->nwalks) return gw_fail(ctx, GW_ESTATE, "gw_summary_read: no summary recorded for that many walks");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (hash) GW_HIP(ctx, hipMemcpy(hash, ctx->d_hash, n * 8, hipMemcpyDeviceToHost));
    if (visits) GW_HIP(ctx, hipMemcpy(visits, ctx->d_visits, n * 4, hipMemcpyDeviceToHost));
    return GW_OK;
}

int gw_sources(uint64_t seed, uint32_t n, uint32_t count, uint32_t* out) {
    if (!out || n == 0) return GW_EINVAL;
    for (uint32_t k = 0; k < count; ++k) {
        const gw::u32x4 r = gw::philox4x32_10(k, 0, gw::kTagSrcs, 0, static_cast<uint32_t>(seed),
                                              static_cast<uint32_t>(seed >> 32));
        out[k] = gw::umulhi32(r.x, n);
    }
    return GW_OK;
}

int gw_events_read(gw_ctx* ctx, uint64_t* out, uint64_t n_walks, uint32_t cap) {
    if (!ctx || !out) return gw_fail(ctx, GW_EINVAL, "gw_events_read: bad arguments");
    if (ctx->pool) {
        const gw_pool& P = *ctx->pool;
        if (!P.events_valid || cap != P.h_events_cap || n_walks > P.n)
            return gw_fail(ctx, GW_ESTATE, "gw_events_read: no event log of that shape");
        std::copy(P.h_events, P.h_events + n_walks * cap, out);
        return GW_OK;
    }
    if (!ctx->d_events || cap != ctx->events_cap || n_walks > ctx->walk_cap)
        return gw_fail(ctx, GW_ESTATE, "gw_events_read: no event log of that shape");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    GW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (n_walks) GW_HIP(ctx, hipMemcpy(out, ctx->d_events, n_walks * cap * 8ull, hipMemcpyDeviceToHost));
    return GW_OK;
}

int gw_gather_probe(gw_ctx* ctx, uint64_t bytes, uint64_t gathers, double* ms) {
    if (!ctx || bytes < 4096 || gathers == 0 || bytes / 4 > 0xFFFFFFFFull) return gw_fail(ctx, GW_EINVAL, "gw_gather_probe: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t *buf = nullptr, *out = nullptr;
    const uint32_t tb = 256, grid = std::max(1u, ctx->cu_count) * 64;
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&buf, bytes, "gather buffer"))) return rc;
    if ((rc = gw_alloc(ctx, (void**)&out, static_cast<uint64_t>(grid) * tb * 4, "gather out"))) { gw_free(ctx, buf, bytes); return rc; }
    hipError_t e = hipMemsetAsync(buf, 1, bytes, ctx->stream);
    const uint32_t nwords = static_cast<uint32_t>(bytes / 4);
    if (e == hipSuccess) {
        gather_probe_kernel<<<dim3(grid), dim3(tb), 0, ctx->stream>>>(buf, nwords, gathers / 8, out);  // warm
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[0], ctx->stream);
    if (e == hipSuccess) {
        gather_probe_kernel<<<dim3(grid), dim3(tb), 0, ctx->stream>>>(buf, nwords, gathers, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[1]);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]);
    gw_free(ctx, buf, bytes);
    gw_free(ctx, out, static_cast<uint64_t>(grid) * tb * 4);
    if (e != hipSuccess) return gw_hip_check(ctx, e, "gather probe");
    if (ms) *ms = t;
    return GW_OK;
}

}  // extern "C"
