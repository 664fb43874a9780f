// gw_internal.h -- engine context and the launchers shared by the .hip units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gwhip.h"

// Source fields below this bound get a partitioned key log (one region per field).
constexpr uint32_t kPartFieldLimit = 1u << 20;
// GW_SRC_SORT picks the partitioned aggregation up to this mean region size (keys per source field).
constexpr uint64_t kPartAutoRegionKeys = 1ull << 17;

// Device counters (unsigned long long each).
constexpr uint32_t kCounters = 32;
constexpr uint32_t kCntSteps = 0, kCntVisits = 1, kCntWorkCursor = 2;
constexpr uint32_t kCntBucketTotal = 8;       // bucketing scan total
constexpr uint32_t kCntExtOverflow = 9;       // successor-header build: degree >= 2^32
constexpr uint32_t kCntTableOverflow = 10;    // hash source counts: table full
constexpr uint32_t kCntCompact = 11;          // hash read-out: occupied slots
constexpr uint32_t kCntLogCursor = 12;        // key log slots claimed
constexpr uint32_t kCntLogOverflow = 13;      // key log full
constexpr uint32_t kCntRuns = 14;             // global sort-reduce: runs
constexpr uint32_t kCntPartOverflow = 15;     // partitioned: keys that missed their region
constexpr uint32_t kCntOutPairs = 16;         // partitioned: exact pairs emitted
constexpr uint32_t kCntSplitPairs = 17;       // partitioned: partial pairs of flushed sources
constexpr uint32_t kCntFlushed = 18;          // partitioned: flushed sources
constexpr uint32_t kCntSplitRuns = 19;        // partitioned: sums of the split pairs
constexpr uint32_t kCntBadWalks = 21;         // persistent schedule: walks skipped at pickup (vertex or restart source not in the graph)
constexpr uint32_t kCntAggItems = 22;         // partitioned: sources with region keys (aggregation work list)

struct gw_run_graph;  // gw_step.hip: a captured run (gw_run_capture)

// Host walk pool (gw_walks_load_host): the records stay in host memory, walk
// id == position, and a run steps them window by window through two device
// slots.  Slot 0 is the context's own walk arrays (sized to the window); slot
// 1 is held here.  Per-walk outputs are copied back by walk id.
struct gw_pool {
    const gw_walk* h_walks = nullptr;
    uint64_t n = 0;
    uint64_t window = 0;
    bool keep_final = false;
    gw_walk* d_in1 = nullptr;
    gw_walk* d_final1 = nullptr;
    uint64_t* d_hash1 = nullptr;
    uint32_t* d_visits1 = nullptr;
    uint64_t* d_events1 = nullptr;
    uint32_t events1_cap = 0;
    gw_walk* h_final = nullptr;    // pinned, n records (keep_final)
    uint64_t* h_hash = nullptr;    // pinned, n
    uint32_t* h_visits = nullptr;  // pinned, n
    uint64_t* h_events = nullptr;  // pinned, n x h_events_cap
    uint32_t h_events_cap = 0;
    bool final_valid = false, summary_valid = false, events_valid = false;
    unsigned long long* h_cnt = nullptr;  // pinned, 2 x kCounters: each slot's counters
    hipStream_t out_stream = nullptr;     // device-to-host copies (the copy stream takes host-to-device)
    hipEvent_t in_ev[2] = {nullptr, nullptr};
    hipEvent_t out_ev[2] = {nullptr, nullptr};
    hipEvent_t run_ev[2] = {nullptr, nullptr};
};

struct gw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cu_count = 0;
    uint64_t hbm_budget = 0;
    uint64_t hbm_used = 0;
    std::string err;

    // graph (resident in HBM)
    uint32_t n = 0;
    uint64_t m = 0;
    uint64_t graph_epoch = 0;  // bumped whenever the resident graph is freed or rewritten
    uint64_t* d_off = nullptr;
    uint32_t* d_col = nullptr;
    uint32_t nblocks = 0;
    std::vector<uint32_t> block_first;  // host copy, nblocks + 1
    uint32_t* d_block_first = nullptr;
    void* d_ext = nullptr;              // successor-header edges (m x 16 B), built on first use
    bool ext_failed = false;

    // host-resident (out-of-core) graph: the CSR stays in pinned host memory
    // and blocks are copied into a device block cache as rounds need them
    bool host_graph = false;
    const uint64_t* h_off = nullptr;
    const uint32_t* h_col = nullptr;
    void* h_owned_off = nullptr;        // pinned copies the context made of caller arrays
    void* h_owned_col = nullptr;
    uint64_t cache_bytes = 0;
    std::vector<std::pair<void*, uint64_t>> pinned;  // gw_host_alloc allocations
    std::vector<int32_t> slot_block;    // block held by each cache slot, -1 = empty
    std::vector<int32_t> block_slot;    // cache slot of each block, -1 = not resident
    std::vector<void*> slot_mem;        // per slot: offsets then columns
    uint64_t slot_bytes = 0;
    void* h_views = nullptr;            // pinned: per block {rebased offsets, rebased columns}
    void* d_views = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_ev[3] = {nullptr, nullptr, nullptr};
    // which blocks a load pass brings in (gw_stream_policy)
    uint32_t block_policy = 0;          // GW_BLOCKS_*
    double block_prob = 0.2;            // GW_BLOCKS_MIN_STEP: chance of the min-step choice
    uint64_t block_seed = 0;            // its coin stream
    uint32_t* d_block_minhop = nullptr; // per-block minimum hop from the bucketing pass
    uint32_t block_minhop_cap = 0;

    // walks
    uint64_t nwalks = 0;
    uint64_t walk_cap = 0;
    uint32_t walk_src_max = 0;  // largest source field among loaded walks
    uint32_t walk_v_max = 0;    // largest start vertex among loaded walks (checked against the graph per run)
    gw_pool* pool = nullptr;    // host walk pool (gw_walks_load_host), else null
    void* serve_state = nullptr; // the open server (gw_serve.hip), else null
    uint32_t serve_page_log2 = 0;   // gw_serve_cache: column page size of a server over a host graph (0 = auto)
    uint32_t serve_max_loads = 0;   // gw_serve_cache: page loads in flight at once (0 = 4)
    uint32_t* d_meet = nullptr; // gw_meeting_counts output, kept between calls
    uint64_t meet_bytes = 0;
    gw_walk* d_walks = nullptr;
    gw_walk* d_walks_alt = nullptr;
    gw_walk* d_final = nullptr;      // final record of each walk, indexed by walk id
    bool finals_valid = true;        // false after a GW_REC_NO_FINAL run
    gw_walk* d_work = nullptr;       // blocked schedule: second bucket buffer
    // Walks placed by gw_walks_init_sources are a function of their index
    // (walk i: id i, source entry i / gen_wps, hop 0), so the persistent
    // kernel can start them from one entry per source (SourceStart, built per
    // run into d_gen_table) instead of a record per walk and a row read per
    // start.  gen_wps = 0: the walk array holds other walks.
    uint32_t* d_gen_src = nullptr;    // source vertices [gen_cap]
    uint32_t* d_gen_field = nullptr;  // source fields [gen_cap] (the vertices when none are given)
    void* d_gen_table = nullptr;      // SourceStart[gen_cap]
    uint32_t gen_cap = 0, gen_nsrc = 0, gen_wps = 0;
    uint64_t gen_built[3] = {~0ull, 0, 0};  // {graph_epoch, field-vertex array, fields} the table was built for

    // recorder outputs
    uint32_t* d_counts = nullptr;   // n
    uint64_t* d_hash = nullptr;     // walk_cap, indexed by walk id
    uint32_t* d_visits = nullptr;   // walk_cap, indexed by walk id
    uint64_t* d_events = nullptr;   // walk_cap x events_cap event slots
    uint32_t events_cap = 0;
    unsigned long long* d_table = nullptr;  // (source, vertex) -> count hash slots
    uint64_t table_slots = 0, table_bytes = 0;
    uint32_t table_log2 = 0, table_vbits = 0;
    uint32_t table_cbits = 0;                // 64 - key bits (count bits of a packed entry)
    bool table_wide = false;                 // 16-byte entries {key + 1, count} for keys over 44 bits
    uint32_t src_method = 1;                 // GW_SRC_* of the last source-count run
    unsigned long long* d_keys = nullptr;    // sort-reduce key log
    uint64_t keys_cap = 0;
    uint64_t keys_sentinel = 0;
    unsigned long long* d_keys_alt = nullptr;  // sort output
    void* d_sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    unsigned long long* d_runs = nullptr;    // (source, vertex) keys of the last source-count run
    uint32_t* d_run_counts = nullptr;
    uint64_t runs_cap = 0, n_runs = 0;
    // Partitioned source counts.  Walks per source-field value of the loaded
    // walks (empty when a field is kPartFieldLimit or more) size one region
    // of the partitioned key log per field.
    std::vector<uint64_t> field_walks;
    // per source field: a designated vertex (the first loaded walk's start)
    // whose visits a direct-region run counts apart; 0xFFFFFFFF for none
    std::vector<uint32_t> field_vertex;
    uint32_t part_per_walk = 0;              // visits per walk the uploaded regions are sized for (0 = stale)
    uint32_t part_fields = 0, part_fields_cap = 0;
    uint64_t part_total = 0;                 // region slots in all (sum of the caps)
    uint32_t* d_part_keys = nullptr;         // regions of vertex ids; after aggregation an exact source's
                                             // region starts with its distinct vertices
    uint32_t* d_part_counts = nullptr;       // visit count beside each distinct vertex
    uint64_t part_keys_cap = 0;
    uint64_t* d_part_off = nullptr;          // region start per field (fields + 1)
    unsigned long long* d_part_cursor = nullptr;  // keys written per field
    unsigned long long* d_part_pairs = nullptr;   // distinct vertices per field (0 for a flushed source)
    uint32_t* d_part_fvert = nullptr;             // field_vertex on the device
    uint32_t* d_part_fcount = nullptr;            // visits of each field's designated vertex
    void* d_part_items = nullptr;                 // aggregation work list, 32 B per field
    uint32_t* d_split_counts = nullptr;      // 2 x split_cap: counts of the split pairs (their keys go to the key log)
    uint64_t split_cap = 0;
    uint64_t part_pairs = 0;                 // last run: pairs held in the regions
    uint64_t split_runs = 0;                 // last run: summed split pairs, in d_runs / d_run_counts
    uint64_t part_flushed = 0;               // last run: sources that went through the split sort

    // device counters (kCounters of them, see the kCnt* indices)
    unsigned long long* d_counters = nullptr;
    unsigned long long* h_run_cnt = nullptr;  // pinned: a run's counters read back (kCounters)
    // bucketing scratch
    uint32_t* d_tile_hist = nullptr;
    uint64_t tile_hist_cap = 0;
    unsigned long long* d_bucket_off = nullptr;  // nblocks + 2

    hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};

    gw_run_graph* run_graph = nullptr;  // the last gw_run_capture

    // gw_run_noalloc: device address of this device's static scratch array
    // (hipGetSymbolAddress at gw_create)
    unsigned long long* d_scratch = nullptr;
};

// Workgroups for a grid-stride loop over n items: at most 2^20, so the
// launch stays far below the 2^32 work-items a dimension may hold (graphs
// reach 2^32 edges at R-MAT scale 28).
inline uint32_t gw_grid(uint64_t n, uint32_t tb) {
    const uint64_t g = (n + tb - 1) / tb;
    return static_cast<uint32_t>(g == 0 ? 1 : g < (1u << 20) ? g : (1u << 20));
}

inline bool gw_has_graph(const gw_ctx* ctx) { return ctx->d_off != nullptr || ctx->host_graph; }
// Frees the graph (device-resident or host-resident) and its block table.
void gw_graph_free(gw_ctx* ctx);
// Walk buffers for n walks (nwalks = n), reallocated only when they grow.
int gw_reserve_walks(gw_ctx* ctx, uint64_t n);

// Record an error message on the context and return `code`.
int gw_fail(gw_ctx* ctx, int code, const std::string& msg);
int gw_hip_check(gw_ctx* ctx, hipError_t e, const char* what);
int gw_alloc(gw_ctx* ctx, void** p, uint64_t bytes, const char* what);
void gw_free(gw_ctx* ctx, void* p, uint64_t bytes);

#define GW_HIP(ctx, call)                                        \
    do {                                                         \
        hipError_t _e = (call);                                  \
        if (_e != hipSuccess) return gw_hip_check(ctx, _e, #call); \
    } while (0)

// gw_rmat.hip
int gw_rmat_build(gw_ctx* ctx, uint32_t scale, uint64_t m, uint64_t seed, uint32_t ta, uint32_t tab,
                  uint32_t tabc);

// gw_rmat.hip: the same R-MAT graph built range by range into pinned host
// memory the context owns, with at most work_bytes of device memory (0 =
// most of what is free); the graph becomes host-resident.
int gw_rmat_build_host(gw_ctx* ctx, uint32_t scale, uint64_t m, uint64_t seed, uint32_t ta, uint32_t tab,
                       uint32_t tabc, uint64_t cache_bytes, uint64_t work_bytes, double* build_ms);

// gw_rmat.hip: CSR on the device from a host edge list (GW_EDGES_* flags)
int gw_edges_build(gw_ctx* ctx, const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t n, uint32_t flags);

// gw_rmat.hip: reverse every edge of the device-resident CSR in place
int gw_transpose_build(gw_ctx* ctx);

// gw_pool.hip: host walk pool.  prepare allocates the slot-1 device buffers
// and the pinned host outputs a run with recorder flags `rec` needs (the
// event log of event_cap entries per walk), and the copy stream and events.
void gw_pool_free(gw_ctx* ctx);
void gw_gen_clear(gw_ctx* ctx);  // the walks are no longer those of gw_walks_init_sources
void gw_gen_free(gw_ctx* ctx);   // and the source arrays are released
int gw_pool_prepare(gw_ctx* ctx, uint32_t rec, uint32_t event_cap);

// gw_step.hip
int gw_step_run(gw_ctx* ctx, const gw_program* prog, gw_stats* stats);
// gw_step.hip: looks up (once per device) the device address of the static
// scratch array gw_run_noalloc uses and keeps it in ctx->d_scratch.
int gw_noalloc_init(gw_ctx* ctx);
void gw_run_graph_free(gw_ctx* ctx);
// Builds the successor-header edges if they fit; false when the CSR is the only layout.
bool gw_ensure_ext(gw_ctx* ctx);

// gw_table.hip: sort + run-length encode the key log (n keys); ms = device time
int gw_keylog_reduce(gw_ctx* ctx, uint64_t n, double* ms);
// gw_table.hip, partitioned source counts.  prepare: size the per-source
// regions for `per_walk` visits per walk and the pair buffers; *ok = false
// when the fields or the memory do not allow it.  reset: zero the device
// counters (and the region cursors when `part`) before a run.  enqueue:
// partition the key log (when `partition`; otherwise the step kernel wrote
// the regions directly) and aggregate every source, with no host sync.
// finish: given the counters read back after the run, grow-and-repeat or
// sort-reduce split pairs as needed; *done = false when a region overflowed
// (the caller falls back to gw_keylog_reduce).  ms = extra device time.
int gw_partlog_prepare(gw_ctx* ctx, uint32_t per_walk, bool* ok);
int gw_partlog_reset(gw_ctx* ctx, bool part);
int gw_partlog_enqueue(gw_ctx* ctx, bool partition);
int gw_partlog_finish(gw_ctx* ctx, const unsigned long long* cnt, bool* done, double* ms);

// gw_stream.hip: device block cache of a host-resident graph.  prepare
// sizes the slots for the current block table (once); reset drops them.
int gw_stream_prepare(gw_ctx* ctx);
void gw_stream_reset(gw_ctx* ctx);
void gw_host_graph_free(gw_ctx* ctx);
// The CSR already in ctx->h_off / ctx->h_col becomes the context's
// host-resident graph (vertex counts, default block table).
int gw_host_graph_adopt(gw_ctx* ctx, uint32_t n, uint64_t m, uint64_t cache_bytes);
// true when the pointer range lies inside one gw_host_alloc allocation
bool gw_is_pinned(const gw_ctx* ctx, const void* p, uint64_t bytes);

// gw_bucket.hip: stable counting sort of `in` (n records) by block into `out`;
// bucket_off[b] = first index of bucket b (nblocks + 1 entries written, host).
// minhop (optional): per-block minimum hop of the records (0xFFFFFFFF: none).
int gw_bucket_sort(gw_ctx* ctx, const gw_walk* in, uint64_t n, gw_walk* out,
                   std::vector<uint64_t>& bucket_off, std::vector<uint32_t>* minhop = nullptr);
