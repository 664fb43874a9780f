// gw_rmat.hip -- R-MAT graph generation and CSR build on the device.
//
// Produces the CSR layout the reference engine reads (uint64 beg_pos,
// uint32 csr; conversions.hpp:181-298), generated directly in HBM:
//   1. edge e -> (src, dst) from Philox(seed, e): one 16-bit uniform per level,
//      quadrant a/b/c/d by cumulative 16-bit thresholds (self-loops and
//      duplicates kept, no permutation);
//   2. keys (src << scale | dst) radix-sorted (rocPRIM: a library sort, off the
//      walk path; every sort starts at bit 0);
//   3. col[e] = low `scale` bits of key e; off[v] = first e with src(e) >= v, written
//      by the edge that starts each run (no atomics, deterministic).
// gw_graph_transpose reverses every edge of the resident CSR with the same
// sort + build steps (SimRank walks the in-links).
//
// gw_graph_rmat_host builds the same graph, bit for bit, into pinned host
// memory when it does not fit the device: a histogram pass over every edge
// by the top source bits, a cut of those bins into source ranges whose edges
// fit the build's device memory, then per range a pass that regenerates
// every edge and keeps the range's (Philox makes regenerating free of
// state), a device sort and a CSR fragment copied to its place on the host
// while the next range is generated.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "gw_device.h"
#include "gw_internal.h"

namespace {

// R-MAT levels [lvl0, lvl1) of edge e appended to (s, d); lvl0 is a
// multiple of 8 (Philox call lvl / 8 supplies levels 8j .. 8j + 7, low
// half-word first), so a caller can stop after the top levels and continue.
struct RmatParams {
    uint32_t k0, k1, ta, tab, tabc;
};

__device__ __forceinline__ void rmat_levels(uint64_t e, uint32_t lvl0, uint32_t lvl1, const RmatParams& rp,
                                            uint32_t& s, uint32_t& d) {
    gw::u32x4 r{0, 0, 0, 0};
    for (uint32_t lvl = lvl0; lvl < lvl1; ++lvl) {
        if ((lvl & 7u) == 0)
            r = gw::philox4x32_10(static_cast<uint32_t>(e), static_cast<uint32_t>(e >> 32), gw::kTagRmat, lvl >> 3, rp.k0, rp.k1);
        const uint32_t word = ((lvl & 7u) >> 1) == 0 ? r.x : ((lvl & 7u) >> 1) == 1 ? r.y : ((lvl & 7u) >> 1) == 2 ? r.z : r.w;
        const uint32_t u = (lvl & 1u) ? (word >> 16) : (word & 0xFFFFu);
        const uint32_t sb = u >= rp.tab ? 1u : 0u;
        const uint32_t db = ((u >= rp.ta && u < rp.tab) || u >= rp.tabc) ? 1u : 0u;
        s = (s << 1) | sb;
        d = (d << 1) | db;
    }
}

__global__ void rmat_keys_kernel(uint64_t* __restrict__ keys, uint64_t m, uint32_t scale, RmatParams rp) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < m; e += stride) {
        uint32_t s = 0, d = 0;
        rmat_levels(e, 0, scale, rp, s, d);
        keys[e] = (static_cast<uint64_t>(s) << scale) | d;
    }
}

// Edges per bin of the top `hb` source bits (levels 0 .. hb - 1, hb <= 16):
// a workgroup counts in LDS, then adds its non-zero bins to the global
// histogram.
constexpr uint32_t kHostBinBits = 12;
__global__ __launch_bounds__(256) void rmat_hist_kernel(uint64_t m, uint32_t hb, RmatParams rp,
                                                      unsigned long long* __restrict__ hist) {
    __shared__ uint32_t h[1u << kHostBinBits];
    const uint32_t nb = 1u << hb;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < m; e += stride) {
        uint32_t s = 0, d = 0;
        rmat_levels(e, 0, hb, rp, s, d);
        atomicAdd(&h[s], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], static_cast<unsigned long long>(h[i]));
}

// Keys ((s - vlo) << scale) | d of the edges whose top-hb source bin lies in
// [blo, bhi), appended in any order (the sort that follows orders them):
// one cursor atomic per wavefront.  Keys past `cap` are counted, not written.
__global__ __launch_bounds__(256) void rmat_range_keys_kernel(uint64_t m, uint32_t scale, uint32_t hb, uint32_t blo,
                                                            uint32_t bhi, uint32_t vlo, RmatParams rp,
                                                            uint64_t* __restrict__ keys, uint64_t cap,
                                                            unsigned long long* __restrict__ cursor) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t first = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    // every lane of a wavefront runs the same number of iterations (the
    // ballot below needs them all): the loop bound is rounded up per wavefront
    const uint64_t wave0 = first & ~63ull;
    for (uint64_t e0 = wave0; e0 < m; e0 += stride) {
        const uint64_t e = e0 + (first - wave0);
        bool keep = false;
        uint64_t key = 0;
        if (e < m) {
            uint32_t s = 0, d = 0;
            const uint32_t top = (hb + 7u) & ~7u;  // the top bins come from whole Philox calls
            rmat_levels(e, 0, top < scale ? top : scale, rp, s, d);
            const uint32_t done = top < scale ? top : scale;
            const uint32_t bin = s >> (done - hb);
            if (bin >= blo && bin < bhi) {
                rmat_levels(e, done, scale, rp, s, d);
                keep = true;
                key = (static_cast<uint64_t>(s - vlo) << scale) | d;
            }
        }
        const uint64_t mask = __ballot(keep);
        if (mask) {
            const uint32_t lane = __lane_id();
            const uint32_t leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
            unsigned long long at = 0;
            if (lane == leader) at = atomicAdd(cursor, static_cast<unsigned long long>(__popcll(mask)));
            at = gw::lane_value(at, leader);
            const uint64_t pos = at + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
            if (keep && pos < cap) keys[pos] = key;
        }
    }
}

// One CSR fragment from a range's sorted keys ((s - vlo) << scale) | d,
// cnt of them, local sources below nv: col[i] = d of key i; off[v] = e0 +
// first key index with local source >= v, for every local v < nv.
__global__ void csr_fragment_kernel(const uint64_t* __restrict__ keys, uint64_t cnt, uint32_t scale, uint32_t nv,
                                    uint64_t e0, uint64_t* __restrict__ off, uint32_t* __restrict__ col) {
    const uint64_t dmask = (1ull << scale) - 1;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e <= cnt; e += stride) {
        const int64_t ls = e < cnt ? static_cast<int64_t>(keys[e] >> scale) : static_cast<int64_t>(nv);
        const int64_t lp = e > 0 ? static_cast<int64_t>(keys[e - 1] >> scale) : -1;
        if (e < cnt) col[e] = static_cast<uint32_t>(keys[e] & dmask);
        for (int64_t v = lp + 1; v <= ls && v < static_cast<int64_t>(nv); ++v) off[v] = e0 + e;
    }
}

// Edge list -> sort keys.  The source is the key (stably sorted, so a row
// keeps the list's order) or, sorting rows too, (src << 32) | dst; a dropped
// self-loop gets the key that sorts past every edge and is counted.
__global__ void edge_keys_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t m,
                                 bool drop_loops, bool sort_rows, uint64_t* __restrict__ keys,
                                 uint32_t* __restrict__ vals, unsigned long long* __restrict__ dropped,
                                 unsigned long long* __restrict__ vmax) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    uint32_t mx = 0, nd = 0;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < m; e += stride) {
        const uint32_t s = src[e], d = dst[e];
        const bool loop = drop_loops && s == d;
        keys[e] = loop ? ~0ull : sort_rows ? (static_cast<uint64_t>(s) << 32) | d : s;
        vals[e] = d;
        nd += loop ? 1u : 0u;
        if (!loop) mx = s > mx ? (d > s ? d : s) : (d > mx ? d : mx);
    }
    const uint32_t tot = gw::wave_sum(nd);
    uint32_t m2 = mx;
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = __shfl_xor(m2, o, 64);
        m2 = y > m2 ? y : m2;
    }
    if (__lane_id() == 0) {
        if (tot) atomicAdd(dropped, static_cast<unsigned long long>(tot));
        atomicMax(vmax, static_cast<unsigned long long>(m2));
    }
}

// CSR from edges sorted by source: off[v] = first edge whose source >= v,
// written by the edge that starts each run (sources are the keys' high or
// only bits).
__global__ void csr_from_sources_kernel(const uint64_t* __restrict__ keys, uint64_t m, uint32_t shift, uint32_t n,
                                        uint64_t* __restrict__ off) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e <= m; e += stride) {
        const uint64_t s = e < m ? (keys[e] >> shift) : static_cast<uint64_t>(n);
        const int64_t prev = e > 0 ? static_cast<int64_t>(keys[e - 1] >> shift) : -1;
        for (int64_t v = prev + 1; v <= static_cast<int64_t>(s); ++v) off[v] = e;
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
    __host__ __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
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
        const RmatParams rp{static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), ta, tab, tabc};
        rmat_keys_kernel<<<dim3(gw_grid(m, tb)), dim3(tb), 0, ctx->stream>>>(k_in, m, scale, rp);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = rocprim::radix_sort_keys(nullptr, temp_bytes, k_in, k_out, m, 0u, 2 * scale, ctx->stream);
        if (e == hipSuccess && (rc = gw_alloc(ctx, &temp, temp_bytes, "sort temp"))) { cleanup(); return rc; }
        if (e == hipSuccess)
            e = rocprim::radix_sort_keys(temp, temp_bytes, k_in, k_out, m, 0u, 2 * scale, ctx->stream);
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
        if (e == hipSuccess) e = rocprim::inclusive_scan(nullptr, need_scan, src, esrc, m, MaxOp(), ctx->stream);
        if (e == hipSuccess)
            e = rocprim::radix_sort_keys(nullptr, need_sort, k_in, k_out, m, 0u, 2 * bits, ctx->stream);
        temp_bytes = std::max(need_scan, need_sort);
        if (e == hipSuccess && (rc = gw_alloc(ctx, &temp, temp_bytes, "transpose temp"))) {
            cleanup();
            return rc;
        }
        size_t t = temp_bytes;
        if (e == hipSuccess) e = rocprim::inclusive_scan(temp, t, src, esrc, m, MaxOp(), ctx->stream);
        if (e == hipSuccess) {
            transpose_keys_kernel<<<dim3(gw_grid(m, tb)), dim3(tb), 0, ctx->stream>>>(
                ctx->d_col, esrc, m, bits, k_in);
            e = hipGetLastError();
        }
        t = temp_bytes;
        if (e == hipSuccess) e = rocprim::radix_sort_keys(temp, t, k_in, k_out, m, 0u, 2 * bits, ctx->stream);
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

// R-MAT straight into pinned host memory (the context owns it), range by
// range: see the file comment.  The result equals gw_rmat_build's bit for
// bit: every range's keys are the graph's keys with sources in the range,
// and a range's sorted keys are the global sorted keys of its rows.
int gw_rmat_build_host(gw_ctx* ctx, uint32_t scale, uint64_t m, uint64_t seed, uint32_t ta, uint32_t tab,
                       uint32_t tabc, uint64_t cache_bytes, uint64_t work_bytes, double* build_ms) {
    const uint32_t n = 1u << scale;
    const uint32_t hb = std::min(scale, kHostBinBits);
    const uint32_t nbins = 1u << hb;
    const uint32_t bin_v = n >> hb;  // vertices per bin
    const RmatParams rp{static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), ta, tab, tabc};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    GW_HIP(ctx, hipEventCreate(&t0));
    GW_HIP(ctx, hipEventCreate(&t1));
    GW_HIP(ctx, hipEventRecord(t0, ctx->stream));
    // 1. pinned CSR arrays
    void* h_off = nullptr;
    void* h_col = nullptr;
    hipError_t e = hipHostMalloc(&h_off, (n + 1ull) * 8, hipHostMallocDefault);
    if (e == hipSuccess && m) e = hipHostMalloc(&h_col, m * 4, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // a refused allocation must not surface at the next launch check
        if (h_off) (void)hipHostFree(h_off);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        return gw_fail(ctx, GW_ENOMEM, "gw_graph_rmat_host: pinned host memory for " + std::to_string(m) +
                                           " edges: " + hipGetErrorString(e));
    }
    uint64_t* hoff = static_cast<uint64_t*>(h_off);
    uint32_t* hcol = static_cast<uint32_t*>(h_col);
    // device buffers, freed on every exit
    unsigned long long* d_hist = nullptr;  // nbins, then the range cursor after them
    uint64_t *k_a = nullptr, *k_b = nullptr;
    uint32_t* c_frag[2] = {nullptr, nullptr};
    uint64_t* o_frag[2] = {nullptr, nullptr};
    void* temp = nullptr;
    size_t temp_bytes = 0;
    uint64_t cap_e = 0, cap_v = 0;
    hipEvent_t copied[2] = {nullptr, nullptr}, ready = nullptr;
    auto cleanup = [&]() {
        if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
        (void)hipStreamSynchronize(ctx->stream);
        gw_free(ctx, d_hist, (nbins + 1ull) * 8);
        gw_free(ctx, k_a, cap_e * 8);
        gw_free(ctx, k_b, cap_e * 8);
        for (int i = 0; i < 2; ++i) {
            gw_free(ctx, c_frag[i], cap_e * 4);
            gw_free(ctx, o_frag[i], cap_v * 8);
            if (copied[i]) (void)hipEventDestroy(copied[i]);
        }
        if (ready) (void)hipEventDestroy(ready);
        gw_free(ctx, temp, temp_bytes);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    };
    auto fail_free = [&](int rc) {
        cleanup();
        (void)hipHostFree(h_off);
        if (h_col) (void)hipHostFree(h_col);
        return rc;
    };
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&d_hist, (nbins + 1ull) * 8, "rmat host-build histogram"))) return fail_free(rc);
    // 2. edges per top-hb source bin
    std::vector<unsigned long long> hist(nbins, 0);
    const uint32_t tb = 256;
    const uint32_t grid = std::max(1u, ctx->cu_count) * 16u;
    if ((e = hipMemsetAsync(d_hist, 0, (nbins + 1ull) * 8, ctx->stream)) != hipSuccess)
        return fail_free(gw_hip_check(ctx, e, "rmat histogram clear"));
    if (m) {
        rmat_hist_kernel<<<dim3(grid), dim3(tb), 0, ctx->stream>>>(m, hb, rp, d_hist);
        if ((e = hipGetLastError()) != hipSuccess) return fail_free(gw_hip_check(ctx, e, "rmat histogram"));
    }
    if ((e = hipMemcpyAsync(hist.data(), d_hist, nbins * 8ull, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return fail_free(gw_hip_check(ctx, e, "rmat histogram"));
    // 3. ranges of whole bins, each within the build's device memory: keys
    //    and their sort buffer (16 B per edge), two fragment buffers of
    //    columns (8 B per edge) and offsets (16 B per vertex)
    size_t free_b = 0, total_b = 0;
    if ((e = hipMemGetInfo(&free_b, &total_b)) != hipSuccess) return fail_free(gw_hip_check(ctx, e, "hipMemGetInfo"));
    uint64_t budget = work_bytes ? work_bytes : static_cast<uint64_t>(free_b) / 10 * 7;
    if (ctx->hbm_budget) budget = std::min<uint64_t>(budget, ctx->hbm_budget > ctx->hbm_used ? ctx->hbm_budget - ctx->hbm_used : 0);
    {  // sort temporaries for the largest range the budget allows
        rocprim::double_buffer<uint64_t> db(nullptr, nullptr);
        const uint64_t most = std::max<uint64_t>(1, std::min<uint64_t>(m, budget / 24));
        if ((e = rocprim::radix_sort_keys(nullptr, temp_bytes, db, most, 0u, 64u, ctx->stream)) != hipSuccess)
            return fail_free(gw_hip_check(ctx, e, "rmat host-build sort size"));
    }
    const uint64_t reserve = temp_bytes + (1ull << 20);
    struct Range { uint32_t blo, bhi; uint64_t edges, e0; };
    std::vector<Range> ranges;
    uint64_t max_e = 0, max_v = 0, e0 = 0;
    for (uint32_t b = 0; b < nbins;) {
        Range r{b, b, 0, e0};
        while (r.bhi < nbins) {
            const uint64_t ne = r.edges + hist[r.bhi];
            const uint64_t nv = static_cast<uint64_t>(r.bhi + 1 - r.blo) * bin_v;
            if (r.bhi > r.blo && ne * 24 + nv * 16 + reserve > budget) break;
            r.edges = ne;
            ++r.bhi;
        }
        if (r.edges * 24 + static_cast<uint64_t>(r.bhi - r.blo) * bin_v * 16 + reserve > budget) {
            return fail_free(gw_fail(ctx, GW_ENOMEM, "gw_graph_rmat_host: source bin " + std::to_string(b) + " holds " +
                                               std::to_string(r.edges) + " edges, more than the build's device memory (" +
                                               std::to_string(budget) + " bytes, " + std::to_string(reserve) +
                                               " of them for the sort's temporaries) can sort at once"));
        }
        max_e = std::max(max_e, r.edges);
        max_v = std::max<uint64_t>(max_v, static_cast<uint64_t>(r.bhi - r.blo) * bin_v);
        e0 += r.edges;
        ranges.push_back(r);
        b = r.bhi;
    }
    cap_e = std::max<uint64_t>(1, max_e);
    cap_v = std::max<uint64_t>(1, max_v);
    if ((rc = gw_alloc(ctx, (void**)&k_a, cap_e * 8, "rmat host-build keys")) ||
        (rc = gw_alloc(ctx, (void**)&k_b, cap_e * 8, "rmat host-build keys (sort)")) ||
        (rc = gw_alloc(ctx, (void**)&c_frag[0], cap_e * 4, "rmat host-build columns")) ||
        (rc = gw_alloc(ctx, (void**)&c_frag[1], cap_e * 4, "rmat host-build columns")) ||
        (rc = gw_alloc(ctx, (void**)&o_frag[0], cap_v * 8, "rmat host-build offsets")) ||
        (rc = gw_alloc(ctx, (void**)&o_frag[1], cap_v * 8, "rmat host-build offsets")))
        return fail_free(rc);
    {
        rocprim::double_buffer<uint64_t> db(k_a, k_b);
        if ((e = rocprim::radix_sort_keys(nullptr, temp_bytes, db, cap_e, 0u, 64u, ctx->stream)) != hipSuccess)
            return fail_free(gw_hip_check(ctx, e, "rmat host-build sort size"));
    }
    if ((rc = gw_alloc(ctx, &temp, temp_bytes, "rmat host-build sort temp"))) return fail_free(rc);
    if (!ctx->copy_stream && (e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking)) != hipSuccess)
        return fail_free(gw_hip_check(ctx, e, "copy stream"));
    for (int i = 0; i < 2; ++i)
        if ((e = hipEventCreateWithFlags(&copied[i], hipEventDisableTiming)) != hipSuccess)
            return fail_free(gw_hip_check(ctx, e, "copy events"));
    if ((e = hipEventCreateWithFlags(&ready, hipEventDisableTiming)) != hipSuccess)
        return fail_free(gw_hip_check(ctx, e, "copy events"));
    // 4. per range: regenerate and keep its edges, sort, CSR fragment, copy
    //    to the host on the copy stream while the next range is generated
    std::vector<unsigned long long> got(ranges.size(), 0);
    unsigned long long* h_got = nullptr;  // pinned cursor read-backs
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_got), ranges.size() * 8, hipHostMallocDefault)) != hipSuccess)
        return fail_free(gw_hip_check(ctx, e, "pinned cursor buffer"));
    unsigned long long* cursor = d_hist + nbins;
    for (size_t i = 0; i < ranges.size(); ++i) {
        const Range& r = ranges[i];
        const int b = static_cast<int>(i & 1);
        const uint32_t vlo = r.blo * bin_v;
        const uint32_t nv = (r.bhi - r.blo) * bin_v;
        uint32_t vbits = 0;
        while (vbits < 32 && (1ull << vbits) < nv) ++vbits;
        if (i >= 2) e = hipStreamWaitEvent(ctx->stream, copied[b], 0);  // fragment buffer b free again
        if (e == hipSuccess) e = hipMemsetAsync(cursor, 0, 8, ctx->stream);
        if (e == hipSuccess && m) {
            rmat_range_keys_kernel<<<dim3(gw_grid(m, tb)), dim3(tb), 0, ctx->stream>>>(m, scale, hb, r.blo, r.bhi, vlo, rp,
                                                                                      k_a, r.edges, cursor);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(h_got + i, cursor, 8, hipMemcpyDeviceToHost, ctx->stream);
        rocprim::double_buffer<uint64_t> db(k_a, k_b);
        if (e == hipSuccess && r.edges) {
            size_t t = temp_bytes;
            e = rocprim::radix_sort_keys(temp, t, db, r.edges, 0u, scale + vbits, ctx->stream);
        }
        if (e == hipSuccess) {
            csr_fragment_kernel<<<dim3(gw_grid(r.edges + 1, tb)), dim3(tb), 0, ctx->stream>>>(
                db.current(), r.edges, scale, nv, r.e0, o_frag[b], c_frag[b]);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(ready, ctx->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->copy_stream, ready, 0);
        if (e == hipSuccess && r.edges)
            e = hipMemcpyAsync(hcol + r.e0, c_frag[b], r.edges * 4, hipMemcpyDeviceToHost, ctx->copy_stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hoff + vlo, o_frag[b], nv * 8ull, hipMemcpyDeviceToHost, ctx->copy_stream);
        if (e == hipSuccess) e = hipEventRecord(copied[b], ctx->copy_stream);
        if (e != hipSuccess) {
            (void)hipHostFree(h_got);
            return fail_free(gw_hip_check(ctx, e, "rmat host-build range"));
        }
    }
    e = hipStreamSynchronize(ctx->copy_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(t1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    for (size_t i = 0; i < ranges.size(); ++i) got[i] = h_got[i];
    (void)hipHostFree(h_got);
    if (e != hipSuccess) return fail_free(gw_hip_check(ctx, e, "rmat host-build"));
    for (size_t i = 0; i < ranges.size(); ++i)
        if (got[i] != ranges[i].edges)
            return fail_free(gw_fail(ctx, GW_EHIP, "gw_graph_rmat_host: range " + std::to_string(i) + " regenerated " +
                                                       std::to_string(got[i]) + " edges, its histogram counted " +
                                                       std::to_string(ranges[i].edges)));
    hoff[n] = m;
    cleanup();
    if (build_ms) *build_ms = ms;
    ctx->h_owned_off = h_off;
    ctx->h_owned_col = h_col;
    ctx->h_off = hoff;
    ctx->h_col = hcol;
    return gw_host_graph_adopt(ctx, n, m, cache_bytes);
}

// A CSR built on the device from a host edge list (the job of the
// reference's convert_to_csr, conversions.hpp:181-298): edges grouped by
// source with a stable radix sort, so a row keeps the list's edge order
// unless rows are sorted too; self-loops dropped on request (the converter
// drops them, conversions.hpp:237); n = the largest vertex + 1 when 0 (the
// converter's vertex count).
int gw_edges_build(gw_ctx* ctx, const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t n, uint32_t flags) {
    const bool drop = (flags & GW_EDGES_DROP_SELF_LOOPS) != 0, sort_rows = (flags & GW_EDGES_SORT_ROWS) != 0;
    int rc;
    uint32_t *d_src = nullptr, *d_dst = nullptr, *v_in = nullptr, *v_out = nullptr;
    uint64_t *k_in = nullptr, *k_out = nullptr;
    unsigned long long* d_cnt = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    const uint64_t mm = std::max<uint64_t>(1, m);
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(ctx->stream);
        gw_free(ctx, d_src, mm * 4);
        gw_free(ctx, d_dst, mm * 4);
        gw_free(ctx, v_in, mm * 4);
        gw_free(ctx, v_out, mm * 4);
        gw_free(ctx, k_in, mm * 8);
        gw_free(ctx, k_out, mm * 8);
        gw_free(ctx, d_cnt, 16);
        gw_free(ctx, temp, temp_bytes);
    };
    if ((rc = gw_alloc(ctx, (void**)&d_src, mm * 4, "edge sources")) ||
        (rc = gw_alloc(ctx, (void**)&d_dst, mm * 4, "edge targets")) ||
        (rc = gw_alloc(ctx, (void**)&v_in, mm * 4, "edge sort values")) ||
        (rc = gw_alloc(ctx, (void**)&v_out, mm * 4, "edge sort values (sorted)")) ||
        (rc = gw_alloc(ctx, (void**)&k_in, mm * 8, "edge sort keys")) ||
        (rc = gw_alloc(ctx, (void**)&k_out, mm * 8, "edge sort keys (sorted)")) ||
        (rc = gw_alloc(ctx, (void**)&d_cnt, 16, "edge counters"))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemsetAsync(d_cnt, 0, 16, ctx->stream);
    if (e == hipSuccess && m) e = hipMemcpyAsync(d_src, src, m * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && m) e = hipMemcpyAsync(d_dst, dst, m * 4, hipMemcpyHostToDevice, ctx->stream);
    const uint32_t tb = 256;
    if (e == hipSuccess && m) {
        edge_keys_kernel<<<dim3(gw_grid(m, tb)), dim3(tb), 0, ctx->stream>>>(d_src, d_dst, m, drop, sort_rows, k_in, v_in,
                                                                             d_cnt, d_cnt + 1);
        e = hipGetLastError();
    }
    unsigned long long cnt[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt, d_cnt, 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        cleanup();
        return gw_hip_check(ctx, e, "edge list upload");
    }
    const uint64_t kept = m - cnt[0];
    const uint64_t nv = n ? n : (kept ? cnt[1] + 1 : 1);
    if (kept && cnt[1] >= nv) {
        cleanup();
        return gw_fail(ctx, GW_EINVAL, "gw_graph_set_edges: vertex " + std::to_string(cnt[1]) + " not below n = " +
                                           std::to_string(nv));
    }
    if (nv >= 0xFFFFFFFFull) {
        cleanup();
        return gw_fail(ctx, GW_EINVAL, "gw_graph_set_edges: vertex ids must stay below 2^32 - 1");
    }
    // the keys sort on every bit (a begin bit above 0 is not used here)
    const uint32_t end_bit = sort_rows ? 64u : (drop && cnt[0] ? 64u : 32u);
    if (m) {
        e = rocprim::radix_sort_pairs(nullptr, temp_bytes, k_in, k_out, v_in, v_out, m, 0u, end_bit, ctx->stream);
        if (e == hipSuccess && (rc = gw_alloc(ctx, &temp, temp_bytes, "edge sort temp"))) {
            cleanup();
            return rc;
        }
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(temp, temp_bytes, k_in, k_out, v_in, v_out, m, 0u, end_bit, ctx->stream);
        if (e != hipSuccess) {
            cleanup();
            return gw_hip_check(ctx, e, "edge sort");
        }
    }
    gw_free(ctx, k_in, mm * 8);
    k_in = nullptr;
    gw_free(ctx, d_src, mm * 4);
    d_src = nullptr;
    if ((rc = gw_alloc(ctx, (void**)&ctx->d_off, (nv + 1) * 8, "offsets")) ||
        (rc = gw_alloc(ctx, (void**)&ctx->d_col, std::max<uint64_t>(1, kept) * 4, "columns"))) {
        cleanup();
        return rc;
    }
    ctx->n = static_cast<uint32_t>(nv);
    ctx->m = kept;
    csr_from_sources_kernel<<<dim3(gw_grid(kept + 1, tb)), dim3(tb), 0, ctx->stream>>>(k_out, kept, sort_rows ? 32u : 0u,
                                                                                        ctx->n, ctx->d_off);
    e = hipGetLastError();
    if (e == hipSuccess && kept) e = hipMemcpyAsync(ctx->d_col, v_out, kept * 4, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return gw_hip_check(ctx, e, "edge CSR build");
    return GW_OK;
}
