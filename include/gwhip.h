/*
 * gwhip.h -- C ABI of the MI355X walk engine (libgwhip.so).
 *
 * This is the only layer that calls HIP.  The C++ front end (the
 * graphwalker_engine / WalkManager / RandomWalk headers under
 * graphwalker_amd/cxx/) and the Python host module (graphwalker_amd/__init__.py)
 * both sit on top of it.  Plain C types only: pointers, sizes, PODs.
 * No exceptions cross this boundary: every entry point returns an int status
 * (GW_OK = 0, negative = error class) and gw_last_error() describes the last
 * failure on that context.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * GraphWalker tree):
 *   gw_create / gw_destroy     graphwalker_engine ctor/dtor (src/engine/graphwalker.hpp:74-144)
 *   gw_graph_set               loadSubGraph / findSubGraph: beg_pos + csr made
 *                              resident (graphwalker.hpp:168-234); block table
 *                              load_block_range (graphwalker.hpp:146-166)
 *   gw_graph_set_host          the out-of-core engine itself: the graph stays in
 *                              host memory and blocks are made resident in a
 *                              device block cache of nmblocks-like capacity
 *                              (findSubGraph / loadSubGraph / swapOut,
 *                              graphwalker.hpp:168-251; block choice
 *                              chooseBlock, walk.hpp:204-252)
 *   gw_host_alloc / gw_host_free   pinned host memory for such graphs (the
 *                              reference reads blocks with pread into malloc'd
 *                              buffers, graphwalker.hpp:84-100, 184-213)
 *   gw_graph_rmat              (no counterpart: the reference reads a
 *                              preprocessed CSR, conversions.hpp:181-298; this
 *                              builds the same CSR layout on the device)
 *   gw_graph_set_edges         convert_to_csr (preprocess/conversions.hpp:181-298):
 *                              the CSR of an edge list, built on the device
 *   gw_graph_rmat_host         the same, into pinned host memory for graphs
 *                              larger than the device (the reference's graph
 *                              lives on disk and is read block by block)
 *   gw_walks_load              the initial buckets written by startWalksbyApp
 *                              through WalkManager::moveWalk (walk.hpp:93-101)
 *   gw_stream_policy           WalkManager::chooseBlock and its heuristics
 *                              (walk.hpp:204-252) for out-of-core load passes
 *   gw_walks_load_host         the walk pool files (writeWalks2Disk /
 *                              readWalksfromDisk, walk.hpp:103-157): a pool kept
 *                              in host memory, stepped window by window
 *   gw_walks_init_*            device-side walk placement for the app patterns
 *                              of rawrandomwalks.cpp:27-50 / rwdomination.cpp:41-61
 *                              (one walk per vertex) and msppr.cpp:41-76
 *                              (walkspersource walks per source)
 *   gw_run_capture / gw_run_replay / gw_autotune   (no counterpart: hipGraph
 *                              replay of a run and a launch-shape search)
 *   gw_run                     graphwalker_engine::run (graphwalker.hpp:272-311)
 *                              with exec_updates (graphwalker.hpp:257-270)
 *                              calling updateByWalk per walk (randomwalk.hpp:39)
 *   gw_bucket                  the re-bucketing of walks by block done by
 *                              moveWalk + getCurrentWalks (walk.hpp:93-137)
 *   gw_counts_read             the per-vertex recorder of updateInfo
 *                              (rwdomination.cpp:63-66 vertex_value[dstId]++)
 *   gw_source_counts_read      msppr's per-source DiscreteDistribution recorder
 *                              (msppr.cpp:71-76, discretedistribution.hpp:44-57), exact
 *   gw_source_topk             DiscreteDistribution::getTop (discretedistribution.hpp:72-95),
 *                              per source, on the device
 *   gw_counts_topk             get_top_vertices (util/toplist.hpp:86-183) over the
 *                              per-vertex counts, on the device
 *   gw_events_read             the per-step updateInfo callback (randomwalk.hpp:32):
 *                              each walk's visits, replayed on the host in walk order
 *   gw_graph_transpose         (no counterpart: the reference's SimRank reads a
 *                              graph the user transposed; this reverses the
 *                              resident CSR on the device)
 *   gw_walks_init_pairs        simrank's placement: R walks from a and R from b,
 *                              walk i at hop i*L (apps/simrank.cpp:30-59)
 *   gw_meeting_counts          simrank's path recorder + computeResult
 *                              (apps/simrank.cpp:60-84): first-meeting counts
 *                              per walk-pair offset, for many pairs at once
 *   gw_walks_read / gw_summary_read   final walk records and per-walk visit
 *                              summaries (no reference counterpart; test access)
 */
#ifndef GWHIP_H
#define GWHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gw_ctx gw_ctx;

/* status codes */
#define GW_OK 0
#define GW_EINVAL (-1)   /* bad argument */
#define GW_ENOMEM (-2)   /* device allocation failed or budget exceeded */
#define GW_EHIP (-3)     /* HIP runtime error */
#define GW_ESTATE (-4)   /* call out of order (no graph, no walks, ...) */

/* walk programs: the updateByWalk variants */
#define GW_JUMP 0          /* RandomWalkwithJump            randomwalkwithjump.hpp:30-62 */
#define GW_STOP 1          /* RandomWalkwithStop            randomwalkwithstop.hpp:20-52 */
#define GW_PROB 2          /* RandomWalkwithProb            randomwalkwithprob.hpp:20-57 */
#define GW_RESTART_JOINT 3 /* RandomWalkwithRestartwithJoint randomwalkwithrestartwithjoint.hpp:29-61 */
#define GW_RESTART 4       /* RandomWalkwithRestart          randomwalkwithrestart.hpp:33-63 */

/* recorders (bit flags): device replacements for the per-step updateInfo */
#define GW_REC_NONE 0u
#define GW_REC_VERTEX_COUNT 1u /* u32 visits per vertex (rwdomination.cpp:65) */
#define GW_REC_WALK_SUMMARY 2u /* per-walk visit count + trajectory hash */
#define GW_REC_EVENTS 4u       /* per-walk event log of (vertex, hop) visits, replayed to updateInfo
                                  on the host (randomwalk.hpp:32); implies GW_REC_WALK_SUMMARY */
#define GW_REC_SOURCE_COUNT 8u /* sparse u32 visit counts per (source field, vertex): the per-source
                                  visit distributions of msppr (msppr.cpp:71-76), exact */
#define GW_REC_NO_FINAL 16u    /* skip writing each walk's final record (the reference keeps none);
                                  gw_walks_read then refuses until a run writes them again */
#define GW_REC_EVENTS_LAST 32u /* with GW_REC_EVENTS: a walk with more visits than event_cap keeps its
                                  latest visit in the last slot of its log.  With event_cap 1 the log
                                  is each walk's last visit: an updateInfo that looks only at the
                                  final hop (graphlet.cpp:46-52, hop == L-1 and dstId == source) is
                                  fed one visit per walk */

/* schedules */
#define GW_SCHED_AUTO 0u
#define GW_SCHED_PLAIN 1u      /* one walk per thread, whole graph resident */
#define GW_SCHED_PERSISTENT 2u /* persistent wavefronts that refill finished lanes */
#define GW_SCHED_BLOCKED 3u    /* device block loop: bucket by block, step block-local */

typedef struct gw_cfg {
    int device;                 /* HIP device ordinal */
    uint32_t reserved;
    uint64_t hbm_budget_bytes;  /* 0 = no cap */
} gw_cfg;

/* One walk record (16 bytes).  The reference packs source:24|offset:26|hop:14
 * into 64 bits (walk.hpp:69-84); the device record keeps the same fields
 * unpacked, plus the stable walk id that keys the random stream. */
typedef struct gw_walk {
    uint32_t id;     /* walk id: Philox counter word 0 */
    uint32_t source; /* source field (walk.hpp:74-76); restart target for RESTART / RESTART_JOINT */
    uint32_t vertex; /* current vertex, global id */
    uint32_t hop;    /* hop field (walk.hpp:82-84) */
} gw_walk;

typedef struct gw_program {
    uint32_t kind;      /* GW_JUMP .. GW_RESTART */
    uint32_t L;         /* hop cap, or the period of RESTART_JOINT */
    uint64_t stop_thr;  /* follow an edge iff deg > 0 and r0 >= stop_thr (p * 2^32) */
    uint32_t nvertices; /* teleport range of GW_JUMP */
    uint32_t stream;    /* Philox counter word 3 */
    uint64_t seed;      /* Philox key */
    uint32_t recorder;  /* GW_REC_* flags */
    uint32_t schedule;  /* GW_SCHED_* */
    uint32_t block_threads;   /* 0 = default workgroup size */
    uint32_t waves_per_simd;  /* persistent schedule occupancy target, 0 = default */
    uint32_t layout;          /* GW_LAYOUT_*: graph layout the step kernel reads */
    uint32_t event_cap;       /* GW_REC_EVENTS: visits stored per walk (more are counted, not stored) */
    uint32_t table_log2;      /* GW_REC_SOURCE_COUNT with GW_SRC_HASH: log2 of hash slots, 0 = auto */
    uint32_t source_method;   /* GW_REC_SOURCE_COUNT: GW_SRC_* */
    uint32_t walk_chunk;      /* walks a wavefront claims per atomic (persistent and blocked
                                 schedules): a multiple of 64 in [64, 8192], 0 = 512 */
} gw_program;

/* How GW_REC_SOURCE_COUNT aggregates (source, vertex) visits.  The sort-based
 * methods append keys to wave-private slabs of a key log during the run. */
#define GW_SRC_SORT 0u         /* GW_SRC_PARTITIONED when the source fields allow it, else GW_SRC_GLOBAL */
#define GW_SRC_HASH 1u         /* device-wide open-addressing hash, one or two atomics per visit */
#define GW_SRC_GLOBAL 2u       /* radix sort of the whole key log + run-length encode */
#define GW_SRC_PARTITIONED 3u  /* one region per source field (fields < 2^20), filled by the step
                                  kernel itself for capped walks (through per-wavefront LDS
                                  buffers) or scattered from the key log for GW_PROB; each region
                                  counted in an LDS hash table; exact sort-reduce for sources
                                  whose distinct vertices outgrow the table */

/* graph layouts the step kernels read (the CSR is always resident) */
#define GW_LAYOUT_AUTO 0u
#define GW_LAYOUT_CSR 1u  /* offsets then column: two dependent random reads per step */
#define GW_LAYOUT_EXT 2u  /* successor-header edges {dst, deg(dst), row(dst)}: one random read per step */

typedef struct gw_stats {
    uint64_t steps;        /* hop increments (device counter) */
    uint64_t visits;       /* updateInfo-equivalent visits (device counter) */
    uint64_t rounds;       /* block rounds (BLOCKED) or 1 */
    uint64_t launches;     /* step-kernel launches */
    double step_ms;        /* sum of step-kernel durations, HIP events on the engine stream */
    double total_ms;       /* whole gw_run on the engine stream (includes bucketing) */
    double bucket_ms;      /* bucketing kernels (BLOCKED) */
    uint32_t schedule;     /* schedule actually used */
    uint32_t layout;       /* layout actually used */
    double reduce_ms;      /* source-count aggregation after the step kernels (sort-based methods) */
    uint32_t source_method;    /* GW_SRC_* actually used (GW_REC_SOURCE_COUNT) */
    uint32_t flushed_sources;  /* GW_SRC_PARTITIONED: sources that went through the exact split sort */
    uint64_t h2d_bytes;        /* host-resident graph: block bytes copied into the device cache */
    uint32_t blocks_loaded;    /* host-resident graph: block loads */
    uint32_t cache_slots;      /* host-resident graph: device block-cache slots */
    double copy_ms;            /* host-resident graph: block-copy time on the copy stream */
    uint64_t pool_windows;     /* host walk pool: windows stepped */
    uint64_t pool_bytes;       /* host walk pool: walk records and per-walk outputs copied */
} gw_stats;

int gw_create(gw_ctx **out, const gw_cfg *cfg);
void gw_destroy(gw_ctx *ctx);
const char *gw_last_error(gw_ctx *ctx);
int gw_device_info(gw_ctx *ctx, char *name, size_t name_len, uint32_t *cu_count, uint64_t *hbm_bytes);

/* Graph: uint64 offsets off[n+1] (absolute), uint32 columns col[m]; host
 * pointers, copied into HBM.  block_first[nblocks+1] partitions the vertices;
 * pass NULL / 0 for a single block. */
int gw_graph_set(gw_ctx *ctx, const uint64_t *off, const uint32_t *col, uint32_t n, uint64_t m);
int gw_graph_rmat(gw_ctx *ctx, uint32_t scale, uint64_t m, uint64_t seed,
                  uint32_t ta, uint32_t tab, uint32_t tabc);
/* CSR built on the device from a host edge list src[e] -> dst[e] (e < m):
 * edges grouped by source with a stable sort, so each row keeps the list's
 * edge order -- the rows convert_to_csr writes from a list grouped by source
 * (conversions.hpp:181-298) -- or with GW_EDGES_SORT_ROWS its columns
 * ascending; GW_EDGES_DROP_SELF_LOOPS drops u -> u edges as the converter
 * does (conversions.hpp:237).  n = 0: the largest vertex id + 1 (the
 * converter's vertex count); otherwise every id must be below n. */
#define GW_EDGES_DROP_SELF_LOOPS 1u
#define GW_EDGES_SORT_ROWS 2u
int gw_graph_set_edges(gw_ctx *ctx, const uint32_t *src, const uint32_t *dst, uint64_t m, uint32_t n,
                       uint32_t flags);
int gw_graph_blocks(gw_ctx *ctx, const uint32_t *block_first, uint32_t nblocks);
/* Reverse every edge of the device-resident graph (rows become in-neighbour
 * lists, ascending, duplicates kept). */
int gw_graph_transpose(gw_ctx *ctx);
int gw_graph_info(gw_ctx *ctx, uint32_t *n, uint64_t *m, uint32_t *nblocks);
int gw_graph_read(gw_ctx *ctx, uint64_t *off, uint32_t *col);

/* Out-of-core graph.  The CSR stays in host memory; gw_run steps it block by
 * block through a device block cache of at most cache_bytes (slots sized for
 * the largest block), copying blocks on a side stream while resident blocks
 * step.  Arrays allocated with gw_host_alloc are used in place (they must
 * outlive the graph); other host arrays are copied once into pinned memory
 * the context owns.  Without a gw_graph_blocks call the vertices are cut into
 * blocks of at most cache_bytes / 4 bytes each.  Runs use the BLOCKED
 * schedule and the CSR layout; GW_REC_SOURCE_COUNT uses GW_SRC_HASH. */
int gw_host_alloc(gw_ctx *ctx, uint64_t bytes, void **out);
int gw_host_free(gw_ctx *ctx, void *p);
int gw_graph_set_host(gw_ctx *ctx, const uint64_t *off, const uint32_t *col, uint32_t n, uint64_t m,
                      uint64_t cache_bytes);
/* The gw_graph_rmat graph, bit for bit, built straight into pinned host
 * memory the context owns (for graphs larger than the device) and made the
 * host-resident graph, as gw_graph_set_host with cache_bytes.  It is built
 * in source ranges whose edges fit `work_bytes` of device memory (0 = most
 * of what is free): every edge is regenerated per range and the range's are
 * sorted on the device, their CSR fragment copied to the host while the next
 * range is generated.  build_ms (may be NULL): wall time on the device.
 * gw_graph_read copies the arrays out. */
int gw_graph_rmat_host(gw_ctx *ctx, uint32_t scale, uint64_t m, uint64_t seed, uint32_t ta, uint32_t tab,
                       uint32_t tabc, uint64_t cache_bytes, uint64_t work_bytes, double *build_ms);

/* Walks: records with ids < n (ids key the random streams). */
int gw_walks_load(gw_ctx *ctx, const gw_walk *walks, uint64_t n);
/* SimRank placement for npairs pairs (a[k], b[k]): walk ids 2kR + i (from
 * a[k]) and 2kR + R + i (from b[k]), i < R, each starting at hop i*L with its
 * start vertex as source field.  Run them with GW_RESTART_JOINT, period L,
 * and GW_REC_EVENTS with event_cap >= L - 1, then call gw_meeting_counts:
 * counts[k*L + l] = walk pairs (i, j) of pair k whose paths first hold the
 * same vertex at offset l, so SimRank(a[k], b[k]) = sum_l counts[k*L+l] c^l / R^2.
 * L - 1 <= 31. */
int gw_walks_init_pairs(gw_ctx *ctx, const uint32_t *a, const uint32_t *b, uint32_t npairs, uint32_t R, uint32_t L);
int gw_meeting_counts(gw_ctx *ctx, uint32_t npairs, uint32_t R, uint32_t L, uint32_t *counts);
/* count walks per vertex: walk i = vertex i / per_vertex, source = vertex, id = i */
/* Which blocks a load pass of an out-of-core run (gw_graph_set_host) brings
 * into the device block cache, after the resident blocks' own walks:
 *   GW_BLOCKS_MAX_WALKS  the blocks with the most walks, resident blocks
 *                        counting double (default)
 *   GW_BLOCKS_MIN_STEP   the reference's chooseBlock(prob) (walk.hpp:245-252):
 *                        with probability `prob` the block holding the walk
 *                        with the fewest hops first, then as MAX_WALKS
 *   GW_BLOCKS_MAX_WEIGHT blockWithMaxWeight (walk.hpp:228-238): walks / min hop
 * The coin stream is keyed by `seed`.  Results never depend on the policy. */
#define GW_BLOCKS_MAX_WALKS 0u
#define GW_BLOCKS_MIN_STEP 1u
#define GW_BLOCKS_MAX_WEIGHT 2u
int gw_stream_policy(gw_ctx *ctx, uint32_t policy, double prob, uint64_t seed);

/* Host walk pool: `walks` (n records, walks[i].id == i) stays in host memory
 * -- gw_host_alloc memory is copied from directly -- and must outlive the
 * pool.  gw_run steps it `window` walks at a time through two device windows,
 * copying the next window in while one steps; per-vertex counts and source
 * counts (GW_SRC_HASH) accumulate over the windows, per-walk outputs (summary,
 * event log, and final records with GW_POOL_KEEP_FINAL) are copied back into
 * pinned host arrays read by gw_walks_read / gw_summary_read /
 * gw_events_read.  Loading device walks (gw_walks_load, gw_walks_init_*)
 * drops the pool. */
#define GW_POOL_KEEP_FINAL 1u
int gw_walks_load_host(gw_ctx *ctx, const gw_walk *walks, uint64_t n, uint64_t window, uint32_t flags);
int gw_walks_init_per_vertex(gw_ctx *ctx, uint32_t per_vertex);
/* walks_per_source walks per source: id = k * walks_per_source + j, source =
 * source_field[k] (or sources[k] when source_field is NULL), vertex = sources[k].
 * Persistent runs start these walks from one entry per source instead of a
 * record per walk, until other walks are placed or gw_bucket reorders them. */
int gw_walks_init_sources(gw_ctx *ctx, const uint32_t *sources, const uint32_t *source_field,
                          uint32_t nsources, uint32_t walks_per_source);
/* Final record of every walk after the last gw_run, indexed by walk id. */
int gw_walks_read(gw_ctx *ctx, gw_walk *out, uint64_t n);
/* The loaded records in their current slot order (after gw_bucket: by block). */
int gw_walks_current(gw_ctx *ctx, gw_walk *out, uint64_t n);

/* Multi-source workload sources: out[k] = mulhi32(r0, n) with
 * r = Philox4x32-10(counter {k, 0, 'SRCS', 0}, key = seed) -- uniform vertices,
 * reproducible from (seed, k).  No context needed. */
int gw_sources(uint64_t seed, uint32_t n, uint32_t count, uint32_t *out);

/* Run every loaded walk to completion. */
int gw_run(gw_ctx *ctx, const gw_program *prog, gw_stats *stats);
/* hipGraph of one run (persistent or plain schedule; vertex counts, walk
 * summaries, events, hash or partitioned source counts): gw_run_capture runs
 * the program once, then captures its launch sequence -- counter reset,
 * recorder clears, step kernel, aggregation, counter read-back -- into a
 * graph; gw_run_replay launches it and finishes the run as gw_run does
 * (total_ms and step_ms: the whole graph).  A replay is refused once device
 * buffers, the graph or the walks changed since the capture. */
int gw_run_capture(gw_ctx *ctx, const gw_program *prog);
int gw_run_replay(gw_ctx *ctx, gw_stats *stats);
/* Autotuner: runs prog at every (block_threads, waves_per_simd) candidate,
 * then at the best of those with walk_chunk in {128, 512, 2048, 8192}
 * (persistent and blocked schedules), `reps` timed runs each after a
 * warm-up, and returns the fastest in *best.  For the blocked schedule on a
 * device-resident graph it first picks the block size -- uniform tables of
 * 2^12 .. 2^24 vertices per block, at most 2048 blocks -- and leaves the
 * fastest table installed (gw_graph_info shows it).  Results depend on none
 * of these. */
int gw_autotune(gw_ctx *ctx, const gw_program *prog, uint32_t reps, gw_program *best, double *best_ms);

/* gw_run without allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  It runs `prog` with the persistent
 * schedule over `nwalks` walk records the caller holds in device memory, on
 * the context's device-resident graph (the CSR, or the successor-header
 * edges when an earlier gw_run built them), and enqueues everything on
 * `stream` (a hipStream_t; NULL = the context's stream) without waiting:
 *   counts  device u32[n], zeroed then filled (GW_REC_VERTEX_COUNT); NULL
 *           with GW_REC_NONE.  No other recorder, and no final records.
 *   totals  device u64[3]: steps, visits, walks skipped because their vertex
 *           (or a restart program's source field) is not a graph vertex.
 * Results equal gw_run's on the same walks (persistent schedule).  Its
 * counters live in a static 256-byte device array per device, shared by
 * every call on that device: calls that overlap on one device race.
 * nwalks above 2^32 - 1 is rejected (GW_EINVAL). */
int gw_run_noalloc(gw_ctx *ctx, const gw_program *prog, const gw_walk *walks, uint64_t nwalks,
                   uint32_t *counts, unsigned long long *totals, void *stream);

/* Stable counting sort of the loaded walks by block of their vertex (out of
 * range vertices are dropped); bucket_counts[nblocks] receives the sizes. */
int gw_bucket(gw_ctx *ctx, uint64_t *bucket_counts);

int gw_counts_read(gw_ctx *ctx, uint32_t *counts, uint32_t n);
int gw_summary_read(gw_ctx *ctx, uint64_t *hash, uint32_t *visits, uint64_t n);
/* Per-(source, vertex) counts of the last run with GW_REC_SOURCE_COUNT, sorted
 * by source then vertex.  With null arrays only *n_out (the number of pairs)
 * is returned; otherwise `cap` must be at least that number. */
int gw_source_counts_read(gw_ctx *ctx, uint32_t *source, uint32_t *vertex, uint32_t *count, uint64_t cap,
                          uint64_t *n_out);
/* Per-source top-k of the last GW_REC_SOURCE_COUNT run (1 <= k <= 1024): each
 * source's k most-visited vertices, count descending, ties to the smaller
 * vertex id; sources ascending.  NULL arrays: *n_out = entry count only. */
int gw_source_topk(gw_ctx *ctx, uint32_t k, uint32_t *source, uint32_t *vertex, uint32_t *count, uint64_t cap,
                   uint64_t *n_out);
/* Top-k vertices of the per-vertex counts of the last GW_REC_VERTEX_COUNT run:
 * count descending, ties to the smaller vertex id.  k is capped at the vertex
 * count; vertex and count hold at least min(k, n) entries; *n_out = entries
 * written. */
int gw_counts_topk(gw_ctx *ctx, uint64_t k, uint32_t *vertex, uint32_t *count, uint64_t *n_out);
/* Event log of the last run with GW_REC_EVENTS: walk id w's visit k is
 * out[w * cap + k] = (hop << 32) | vertex for k < min(visits[w], cap).
 * `n_walks` and `cap` must equal the loaded walk count and the run's event_cap. */
int gw_events_read(gw_ctx *ctx, uint64_t *out, uint64_t n_walks, uint32_t cap);

/* ---- serving: continuous batching of PPR-style queries -------------------
 * Admitted queries share one device pool of walk records; each gw_serve_step
 * advances every live walk by at most `hops_per_step` hops, so queries that
 * arrive while others are in flight join at the next step.  A query's visits
 * are appended to its own key region of a count arena in HBM while its walks
 * run, and counted into its own hash region when its last walk ends; queries
 * are admitted first come first served while walk slots, a query slot and
 * arena space are free (8-byte slots: 2 x min(walks * L, n) rounded up to a
 * power of two for the hash region, plus walks * L / 2 for the keys).  The
 * arena takes arena_bytes of HBM plus half as much again for each region's
 * list of occupied slots (completion reads only those).  The random stream of walk i of query q is Philox(counter {i, hop,
 * 'WALK', q.qid low word}, key {seed low, seed high ^ q.qid high word}), so a
 * query's result does not depend on the other queries in flight.  There is
 * no counterpart in the reference (it has no serving loop). */
typedef struct gw_query {
    uint64_t qid;     /* caller's id: keys the query's random streams */
    uint32_t source;  /* source vertex */
    uint32_t walks;   /* walk count */
} gw_query;

typedef struct gw_serve_stats {
    uint64_t steps, visits;   /* walk-steps / visits made by this serve step */
    uint64_t live_walks;      /* walks still in flight after the step */
    uint32_t admitted, completed, active_queries, queued;
    double step_ms;           /* step kernel time, HIP events on the engine stream */
    /* host-resident graph only (zero otherwise): */
    uint64_t h2d_bytes;       /* column pages whose copy into the page cache started before this step */
    uint32_t pages_loaded;    /* how many */
    uint32_t pages_resident;  /* pages in the cache during the step */
    uint64_t sampled_reads;   /* 1-in-64 sample of the step's column reads */
    uint64_t sampled_resident;/* of those, reads served by the page cache (the rest crossed the host link) */
} gw_serve_stats;

/* Over a host-resident graph (gw_graph_set_host, gw_graph_rmat_host) the
 * server copies the offsets to HBM and keeps the columns in host memory, read
 * through a device page cache of pages of 2^page_log2 edges: the graph's
 * cache_bytes, offsets included, bound it.  Columns of pages not in the cache
 * are read from the pinned host arrays in place; between serve steps the
 * pages read most (a 1-in-64 sample of each step's reads, decayed) are
 * copied into the cache on the copy stream while steps run, replacing the
 * least-read ones.  Results do not depend on which pages are resident.
 * gw_serve_cache (before gw_serve_open) sets the page size (0 = pages of at
 * least 2^10 edges, at most 4096 pages) and the page copies in flight at once
 * (0 = 4).  Host-resident graphs are served in the CSR layout.
 * max_queries: queries in flight at once, 1 .. 65535 (completed queries are
 * processed one workgroup row per query).  gw_serve_submit refuses a query
 * whose count region cannot fit the arena, and an id already queued, in
 * flight or completed but not yet read by gw_serve_result.  A step that fails
 * after its kernel ran leaves the server failed: only gw_serve_close works. */
int gw_serve_cache(gw_ctx *ctx, uint32_t page_log2, uint32_t max_loads);
int gw_serve_open(gw_ctx *ctx, const gw_program *prog, uint64_t pool_walks, uint64_t arena_bytes,
                  uint32_t max_queries, uint32_t hops_per_step);
int gw_serve_submit(gw_ctx *ctx, const gw_query *queries, uint32_t n);
int gw_serve_step(gw_ctx *ctx, gw_serve_stats *stats);
/* Pops the next completed query: returns 0 and its id and pair count, or 1
 * when none is waiting (queries already read by id are skipped). */
int gw_serve_poll(gw_ctx *ctx, uint64_t *qid, uint64_t *npairs);
/* Copies a completed query's (vertex, count) pairs, in no particular order,
 * and releases them. */
int gw_serve_result(gw_ctx *ctx, uint64_t qid, uint32_t *vertex, uint32_t *count, uint64_t cap);
/* The completed queries not yet read, in completion order, at most max_q of
 * them: qid[i] and npairs[i], their pairs back to back in vertex / count (at
 * most cap entries; a query that does not fit stays for the next call, and
 * GW_EINVAL if the first one does not).  *n_out = queries returned, each
 * released.  One call per serve step in place
 * of a gw_serve_poll + gw_serve_result pair per query (no counterpart in the
 * reference). */
int gw_serve_results(gw_ctx *ctx, uint32_t max_q, uint64_t *qid, uint64_t *npairs, uint32_t *vertex, uint32_t *count,
                     uint64_t cap, uint32_t *n_out);
/* Completed queries keep only their k most-visited vertices (1 <= k <= 1024;
 * 0 = every pair, the default), selected on the device; gw_serve_result then
 * returns them count descending, ties to the smaller vertex id. */
int gw_serve_topk(gw_ctx *ctx, uint32_t k);
int gw_serve_close(gw_ctx *ctx);

/* Random-gather ceiling: `gathers` independent random 4-byte loads from a
 * buffer of `bytes` bytes; returns the kernel time in ms. */
int gw_gather_probe(gw_ctx *ctx, uint64_t bytes, uint64_t gathers, double *ms);
/* Streaming copy of `bytes` bytes between two fresh buffers (the known-size
 * kernel hardware byte counters are calibrated with); kernel time in ms. */
int gw_copy_probe(gw_ctx *ctx, uint64_t bytes, double *ms);
/* Random-read ceiling of pinned host memory from a kernel (how a server over
 * a host-resident graph reads the columns of pages not in its cache):
 * `gathers` independent random 4-byte reads over a pinned buffer of `bytes`
 * bytes (at most 16 GiB); kernel time in ms. */
int gw_host_gather_probe(gw_ctx *ctx, uint64_t bytes, uint64_t gathers, double *ms);
/* Host-to-device copy of `bytes` from pinned host memory on the block-copy
 * stream (the path out-of-core block loads take); copy time in ms. */
int gw_h2d_probe(gw_ctx *ctx, uint64_t bytes, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* GWHIP_H */
