// gw_table.hip -- read-out of the (source, vertex) -> count hash table that
// GW_REC_SOURCE_COUNT fills during a run: non-empty slots are compacted,
// radix-sorted (rocPRIM through hipCUB, off the walk path) and
// unpacked into (source, vertex, count) triples sorted by source, vertex.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

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

}  // namespace

extern "C" int gw_source_counts_read(gw_ctx* ctx, uint32_t* source, uint32_t* vertex, uint32_t* count, uint64_t cap,
                                     uint64_t* n_out) {
    if (!ctx || !n_out) return gw_fail(ctx, GW_EINVAL, "gw_source_counts_read: bad arguments");
    if (!ctx->d_table) return gw_fail(ctx, GW_ESTATE, "gw_source_counts_read: no run with GW_REC_SOURCE_COUNT");
    GW_HIP(ctx, hipSetDevice(ctx->device));
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
