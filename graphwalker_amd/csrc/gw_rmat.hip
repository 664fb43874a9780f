// gw_rmat.hip -- R-MAT graph generation and CSR build on the device.
//
// Produces the CSR layout the reference engine reads (uint64 beg_pos,
// uint32 csr; conversions.hpp:181-298), generated directly in HBM:
//   1. edge e -> (src, dst) from Philox(seed, e): one 16-bit uniform per level,
//      quadrant a/b/c/d by cumulative 16-bit thresholds (self-loops and
//      duplicates kept, no permutation);
//   2. keys (src << scale | dst) radix-sorted (rocPRIM through hipCUB: a library
//      sort, off the walk path);
//   3. col[e] = low `scale` bits of key e; off[v] = first e with src(e) >= v, written
//      by the edge that starts each run (no atomics, deterministic).
// gw_graph_transpose reverses every edge of the resident CSR with the same
// sort + build steps (SimRank walks the in-links).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "gw_device.h"
#include "gw_internal.h"

namespace {

__global__ void rmat_keys_kernel(uint64_t* __restrict__ keys, uint64_t m, uint32_t scale, uint32_t k0, uint32_t k1,
                                 uint32_t ta, uint32_t tab, uint32_t tabc) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < m; e += stride) {
        uint32_t s = 0, d = 0;
        gw::u32x4 r{0, 0, 0, 0};
        for (uint32_t lvl = 0; lvl < scale; ++lvl) {
            if ((lvl & 7u) == 0)
                r = gw::philox4x32_10(static_cast<uint32_t>(e), static_cast<uint32_t>(e >> 32), gw::kTagRmat, lvl >> 3, k0, k1);
            const uint32_t word = ((lvl & 7u) >> 1) == 0 ? r.x : ((lvl & 7u) >> 1) == 1 ? r.y : ((lvl & 7u) >> 1) == 2 ? r.z : r.w;
            const uint32_t u = (lvl & 1u) ? (word >> 16) : (word & 0xFFFFu);
            const uint32_t sb = u >= tab ? 1u : 0u;
            const uint32_t db = ((u >= ta && u < tab) || u >= tabc) ? 1u : 0u;
            s = (s << 1) | sb;
            d = (d << 1) | db;
        }
        keys[e] = (static_cast<uint64_t>(s) << scale) | d;
    }
}

// Keys hold (src << bits) | dst, sorted; n vertices.
__global__ void csr_from_sorted_kernel(const uint64_t* __restrict__ keys, uint64_t m, uint32_t scale, uint32_t n,
                                       uint64_t* __restrict__ off, uint32_t* __restrict__ col) {
    const uint64_t dmask = (1ull << scale) - 1;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e <= m; e += stride) {
        // The run of source s starts at e: off[v] = e for every v in (prev_src, s].
        const uint64_t s = e < m ? (keys[e] >> scale) : static_cast<uint64_t>(n);
        const int64_t prev = e > 0 ? static_cast<int64_t>(keys[e - 1] >> scale) : -1;
        if (e < m) col[e] = static_cast<uint32_t>(keys[e] & dmask);
        for (int64_t v = prev + 1; v <= static_cast<int64_t>(s); ++v) off[v] = e;
    }
}

// Row start markers: marker[off[v]] = v for every non-empty row v (rows are
// sorted, so an inclusive max-scan of the markers gives each edge's source).
__global__ void row_marker_kernel(const uint64_t* __restrict__ off, uint32_t n, uint32_t* __restrict__ marker) {
    const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (v >= n) return;
    if (off[v + 1] > off[v]) marker[off[v]] = static_cast<uint32_t>(v);
}

// Reversed-edge keys (dst << bits) | src.
__global__ void transpose_keys_kernel(const uint32_t* __restrict__ col, const uint32_t* __restrict__ src, uint64_t m,
                                      uint32_t bits, uint64_t* __restrict__ keys) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < m; e += stride)
        keys[e] = (static_cast<uint64_t>(col[e]) << bits) | src[e];
}

struct MaxOp {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

}  // namespace

int gw_rmat_build(gw_ctx* ctx, uint32_t scale, uint64_t m, uint64_t seed, uint32_t ta, uint32_t tab, uint32_t tabc) {
    const uint32_t n = 1u << scale;
    if (scale == 32) return gw_fail(ctx, GW_EINVAL, "gw_graph_rmat: scale 32 needs 64-bit vertex ids");
    int rc;
    uint64_t *k_in = nullptr, *k_out = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    if ((rc = gw_alloc(ctx, (void**)&k_in, m * 8, "rmat keys"))) return rc;
    if ((rc = gw_alloc(ctx, (void**)&k_out, m * 8, "rmat keys (sorted)"))) { gw_free(ctx, k_in, m * 8); return rc; }
    auto cleanup = [&]() {
        gw_free(ctx, k_in, m * 8);
        gw_free(ctx, k_out, m * 8);
        gw_free(ctx, temp, temp_bytes);
    };
    const uint32_t tb = 256;
    hipError_t e = hipSuccess;
    if (m) {
        rmat_keys_kernel<<<dim3(gw_grid(m, tb)), dim3(tb), 0, ctx->stream>>>(
            k_in, m, scale, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), ta, tab, tabc);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, k_in, k_out, m, 0, 2 * scale, ctx->stream);
        if (e == hipSuccess && (rc = gw_alloc(ctx, &temp, temp_bytes, "sort temp"))) { cleanup(); return rc; }
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, k_in, k_out, m, 0, 2 * scale, ctx->stream);
    }
    if (e != hipSuccess) { cleanup(); return gw_hip_check(ctx, e, "rmat generate/sort"); }
    // keys sorted in k_out; k_in is free to be reused for nothing: release it
    // before the CSR arrays to keep the peak at 2 key arrays.
    gw_free(ctx, k_in, m * 8);
    k_in = nullptr;
    gw_free(ctx, temp, temp_bytes);
    temp = nullptr;
    temp_bytes = 0;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_off, (n + 1ull) * 8, "offsets"))) { cleanup(); return rc; }
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_col, m * 4, "columns"))) { cleanup(); return rc; }
    ctx->n = n;
    ctx->m = m;
    csr_from_sorted_kernel<<<dim3(gw_grid(m + 1, tb)), dim3(tb), 0, ctx->stream>>>(
        k_out, m, scale, n, ctx->d_off, ctx->d_col);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return gw_hip_check(ctx, e, "csr build");
    return GW_OK;
}

// Reverses every edge of the device-resident CSR: keys (dst << bits) | src
// from the row markers' max-scan, radix sort, CSR build.  Rows of the result
// list in-neighbours in ascending order, with duplicates kept.
int gw_transpose_build(gw_ctx* ctx) {
    const uint32_t n = ctx->n;
    const uint64_t m = ctx->m;
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < n) ++bits;
    int rc;
    uint32_t* src = nullptr;
    uint64_t *k_in = nullptr, *k_out = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    auto cleanup = [&]() {
        gw_free(ctx, src, m * 4);
        gw_free(ctx, k_in, m * 8);
        gw_free(ctx, k_out, m * 8);
        gw_free(ctx, temp, temp_bytes);
    };
    const uint32_t tb = 256;
    hipError_t e = hipSuccess;
    if (m) {
        if ((rc = gw_alloc(ctx, (void**)&src, m * 4, "transpose sources")) ||
            (rc = gw_alloc(ctx, (void**)&k_in, m * 8, "transpose keys")) ||
            (rc = gw_alloc(ctx, (void**)&k_out, m * 8, "transpose keys (sorted)"))) {
            cleanup();
            return rc;
        }
        e = hipMemsetAsync(src, 0, m * 4, ctx->stream);
        if (e == hipSuccess) {
            row_marker_kernel<<<dim3((n + tb - 1) / tb), dim3(tb), 0, ctx->stream>>>(ctx->d_off, n, src);
            e = hipGetLastError();
        }
        // edge sources: max-scan of the markers into the (still unused) sort
        // output, reversed keys into the sort input, then the sort
        uint32_t* esrc = reinterpret_cast<uint32_t*>(k_out);
        size_t need_scan = 0, need_sort = 0;
        if (e == hipSuccess) e = hipcub::DeviceScan::InclusiveScan(nullptr, need_scan, src, esrc, MaxOp(), m, ctx->stream);
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortKeys(nullptr, need_sort, k_in, k_out, m, 0, 2 * bits, ctx->stream);
        temp_bytes = std::max(need_scan, need_sort);
        if (e == hipSuccess && (rc = gw_alloc(ctx, &temp, temp_bytes, "transpose temp"))) {
            cleanup();
            return rc;
        }
        size_t t = temp_bytes;
        if (e == hipSuccess) e = hipcub::DeviceScan::InclusiveScan(temp, t, src, esrc, MaxOp(), m, ctx->stream);
        if (e == hipSuccess) {
            transpose_keys_kernel<<<dim3(gw_grid(m, tb)), dim3(tb), 0, ctx->stream>>>(
                ctx->d_col, esrc, m, bits, k_in);
            e = hipGetLastError();
        }
        t = temp_bytes;
        if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortKeys(temp, t, k_in, k_out, m, 0, 2 * bits, ctx->stream);
    }
    if (e == hipSuccess) {
        csr_from_sorted_kernel<<<dim3(gw_grid(m + 1, tb)), dim3(tb), 0, ctx->stream>>>(
            k_out, m, bits, n, ctx->d_off, ctx->d_col);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return gw_hip_check(ctx, e, "transpose");
    return GW_OK;
}
