// gw_table.hip -- read-out of the (source, vertex) -> count hash table that
// GW_REC_SOURCE_COUNT fills during a run: non-empty slots are compacted,
// radix-sorted (rocPRIM through hipCUB, off the walk path) and
// unpacked into (source, vertex, count) triples sorted by source, vertex.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "gw_device.h"
#include "gw_internal.h"

namespace {

__global__ void compact_kernel(const unsigned long long* __restrict__ tab, uint64_t slots,
                               unsigned long long* __restrict__ out, unsigned long long* __restrict__ n) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const bool full = i < slots && tab[i] != 0ull;
    // one atomic per wavefront: rank the full slots with a ballot
    const uint64_t mask = __ballot(full);
    if (mask == 0) return;
    const uint32_t lane = __lane_id();
    const uint32_t leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(n, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader, 64);
    if (full) {
        const uint32_t r = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
        out[base + r] = tab[i];
    }
}

__global__ void unpack_kernel(const unsigned long long* __restrict__ e, uint64_t n, uint32_t vbits, uint32_t cbits,
                              uint32_t* __restrict__ s, uint32_t* __restrict__ v, uint32_t* __restrict__ c) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const unsigned long long x = e[i];
    const uint64_t key = x >> cbits;
    c[i] = static_cast<uint32_t>(x & ((1ull << cbits) - 1));
    v[i] = static_cast<uint32_t>(key & ((1ull << vbits) - 1));
    s[i] = static_cast<uint32_t>(key >> vbits);
}

__global__ void unpack_runs_kernel(const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ counts,
                                   uint64_t n, uint32_t vbits, uint32_t* __restrict__ s, uint32_t* __restrict__ v,
                                   uint32_t* __restrict__ c) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = keys[i];
    v[i] = static_cast<uint32_t>(k & ((1ull << vbits) - 1));
    s[i] = static_cast<uint32_t>(k >> vbits);
    c[i] = counts[i];
}

}  // namespace

// Sort-reduce of the key log: radix sort of the n logged slots on the key
// bits plus one (the sentinel bit), then run-length encoding; the runs are
// the (source, vertex) pairs in order with their visit counts.  The trailing
// sentinel run (slab padding), if any, is dropped.
int gw_keylog_reduce(gw_ctx* ctx, uint64_t n, double* ms) {
    const int end_bit = static_cast<int>(64 - ctx->table_cbits) + 1;
    if (n > ctx->runs_cap) {
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
    }
    size_t need_sort = 0, need_rle = 0;
    GW_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, need_sort, ctx->d_keys, ctx->d_keys_alt, n, 0, end_bit,
                                                  ctx->stream));
    GW_HIP(ctx, hipcub::DeviceRunLengthEncode::Encode(nullptr, need_rle, ctx->d_keys_alt, ctx->d_runs,
                                                      ctx->d_run_counts, ctx->d_counters + 14, n, ctx->stream));
    const size_t need = std::max(need_sort, need_rle);
    if (need > ctx->sort_temp_bytes) {
        gw_free(ctx, ctx->d_sort_temp, ctx->sort_temp_bytes);
        ctx->d_sort_temp = nullptr;
        ctx->sort_temp_bytes = 0;
        int rc = gw_alloc(ctx, &ctx->d_sort_temp, need, "key sort temp");
        if (rc) return rc;
        ctx->sort_temp_bytes = need;
    }
    size_t tb = ctx->sort_temp_bytes;
    GW_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    if (n) {
        GW_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(ctx->d_sort_temp, tb, ctx->d_keys, ctx->d_keys_alt, n, 0, end_bit,
                                                      ctx->stream));
        tb = ctx->sort_temp_bytes;
        GW_HIP(ctx, hipcub::DeviceRunLengthEncode::Encode(ctx->d_sort_temp, tb, ctx->d_keys_alt, ctx->d_runs,
                                                          ctx->d_run_counts, ctx->d_counters + 14, n, ctx->stream));
    } else {
        GW_HIP(ctx, hipMemsetAsync(ctx->d_counters + 14, 0, 8, ctx->stream));
    }
    GW_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    unsigned long long runs = 0;
    GW_HIP(ctx, hipMemcpyAsync(&runs, ctx->d_counters + 14, 8, hipMemcpyDeviceToHost, ctx->stream));
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

extern "C" int gw_source_counts_read(gw_ctx* ctx, uint32_t* source, uint32_t* vertex, uint32_t* count, uint64_t cap,
                                     uint64_t* n_out) {
    if (!ctx || !n_out) return gw_fail(ctx, GW_EINVAL, "gw_source_counts_read: bad arguments");
    GW_HIP(ctx, hipSetDevice(ctx->device));
    if (ctx->src_method == GW_SRC_SORT) {
        const uint64_t n = ctx->n_runs;
        *n_out = n;
        if (!source && !vertex && !count) return GW_OK;
        if (!source || !vertex || !count || cap < n)
            return gw_fail(ctx, GW_EINVAL, "gw_source_counts_read: output arrays missing or smaller than the pair count");
        if (n == 0) return GW_OK;
        uint32_t* d_out = nullptr;
        int rc = gw_alloc(ctx, (void**)&d_out, n * 12, "runs unpack");
        if (rc) return rc;
        unpack_runs_kernel<<<dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, ctx->stream>>>(
            ctx->d_runs, ctx->d_run_counts, n, ctx->table_vbits, d_out, d_out + n, d_out + 2 * n);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(source, d_out, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(vertex, d_out + n, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(count, d_out + 2 * n, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        gw_free(ctx, d_out, n * 12);
        if (e != hipSuccess) return gw_hip_check(ctx, e, "runs read-out");
        return GW_OK;
    }
    if (!ctx->d_table) return gw_fail(ctx, GW_ESTATE, "gw_source_counts_read: no run with GW_REC_SOURCE_COUNT");
    const uint64_t slots = ctx->table_slots;
    unsigned long long *a = nullptr, *b = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    uint32_t* d_out = nullptr;
    int rc;
    if ((rc = gw_alloc(ctx, (void**)&a, slots * 8, "table compaction"))) return rc;
    auto release = [&]() {
        gw_free(ctx, a, slots * 8);
        gw_free(ctx, b, slots * 8);
        gw_free(ctx, temp, temp_bytes);
        gw_free(ctx, d_out, slots * 12);
    };
    hipError_t e = hipMemsetAsync(ctx->d_counters + 11, 0, 8, ctx->stream);
    if (e == hipSuccess) {
        compact_kernel<<<dim3(static_cast<uint32_t>((slots + 255) / 256)), dim3(256), 0, ctx->stream>>>(
            ctx->d_table, slots, a, ctx->d_counters + 11);
        e = hipGetLastError();
    }
    unsigned long long n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, ctx->d_counters + 11, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { release(); return gw_hip_check(ctx, e, "table compaction"); }
    *n_out = n;
    if (!source && !vertex && !count) { release(); return GW_OK; }
    if (!source || !vertex || !count || cap < n) {
        release();
        return gw_fail(ctx, GW_EINVAL, "gw_source_counts_read: output arrays missing or smaller than the pair count");
    }
    if (n == 0) { release(); return GW_OK; }
    if ((rc = gw_alloc(ctx, (void**)&b, slots * 8, "table sort"))) { release(); return rc; }
    // Sort the whole 64-bit entry: keys are unique per slot, so this orders by
    // key.  (rocPRIM's onesweep sort of 64-bit keys with begin_bit > 0 returns
    // a non-permutation on gfx950 with ROCm 7.2: measured, so it is avoided.)
    const int begin_bit = 0;
    e = hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, a, b, n, begin_bit, 64, ctx->stream);
    if (e == hipSuccess && (rc = gw_alloc(ctx, &temp, temp_bytes, "table sort temp"))) { release(); return rc; }
    if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, a, b, n, begin_bit, 64, ctx->stream);
    if (e == hipSuccess && (rc = gw_alloc(ctx, (void**)&d_out, slots * 12, "table unpack"))) { release(); return rc; }
    if (e == hipSuccess) {
        unpack_kernel<<<dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, ctx->stream>>>(
            b, n, ctx->table_vbits, ctx->table_cbits, d_out, d_out + n, d_out + 2 * n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(source, d_out, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(vertex, d_out + n, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(count, d_out + 2 * n, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    release();
    if (e != hipSuccess) return gw_hip_check(ctx, e, "table read-out");
    return GW_OK;
}
