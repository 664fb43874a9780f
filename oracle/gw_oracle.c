/*
 * gw_oracle.c -- CPU restatement of GraphWalker's walk-update hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * What it restates (reference paths are relative to the GraphWalker tree):
 *   - the per-walk step programs  (src/walks/randomwalkwithjump.hpp:30-62,
 *     randomwalkwithstop.hpp:20-52, randomwalkwithprob.hpp:20-57,
 *     randomwalkwithrestartwithjoint.hpp:29-61, randomwalkwithrestart.hpp:33-63)
 *     with glibc rand_r replaced by Philox4x32-10 keyed by (seed, walk id, hop);
 *   - the block-local loop structure of updateByWalk (walk stays while its
 *     vertex lies in the executing block, then moveWalk to the block of the
 *     vertex: randomwalkwithjump.hpp:42,57-61; walk.hpp:93-101; block lookup
 *     randomwalk.hpp:74-82);
 *   - the bucketing of walk records by block (walk.hpp:93-137) as a stable
 *     counting sort;
 *   - the per-vertex count recorder of rwdomination (apps/rwdomination.cpp:63-66)
 *     without its data race.
 * Generators that feed the path (not in the reference, which reads an edge
 * list from disk; conversions.hpp:181-298 for the CSR layout):
 *   - R-MAT edge generator keyed by Philox(seed, edge index);
 *   - CSR builder: uint64 offsets, uint32 columns, rows sorted ascending.
 *
 * Philox4x32-10 follows Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers:
 * As Easy as 1, 2, 3" (SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
 *
 * Pinned by: tests/test_oracle.py (Random123 known-answer vectors, the rocRAND
 * header engine as a second implementation, closed-form cycle / star / sink
 * graphs, geometric length laws, chi-square neighbour uniformity, FP64
 * finite-horizon expectations, std-sort equality for bucketing).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GWO_TAG_WALK 0x57414C4Bu /* 'WALK' */
#define GWO_TAG_RMAT 0x524D4154u /* 'RMAT' */
#define GWO_TAG_SRCS 0x53524353u /* 'SRCS' */

enum { GWO_JUMP = 0, GWO_STOP = 1, GWO_PROB = 2, GWO_RESTART_JOINT = 3, GWO_RESTART = 4 };
enum { GWO_MOVED = 0, GWO_FINISHED = 1 };

typedef struct {
    uint32_t kind;
    uint32_t L;          /* hop cap (JUMP/STOP/RESTART) or period (RESTART_JOINT) */
    uint64_t stop_thr;   /* the walk continues along an edge iff r0 >= stop_thr; in [0, 2^32] */
    uint32_t nvertices;  /* teleport range for JUMP */
    uint32_t stream;     /* 4th Philox counter word (query slot in serving) */
    uint64_t seed;
} gwo_program;

typedef struct { uint32_t id, source, vertex, hop; } gwo_walk;

/* ---------------------------------------------------------------- Philox */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

static inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void gwo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

/* Batched form for tests: n counters, one key. */
void gwo_philox_batch(const uint32_t *ctr, uint64_t n, const uint32_t key[2], uint32_t *out) {
    for (uint64_t i = 0; i < n; ++i) philox(ctr + 4 * i, key, out + 4 * i);
}

static inline uint32_t mulhi32(uint32_t a, uint32_t b) {
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}
static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
}

/* ---------------------------------------------------------------- R-MAT */
/* Edge e draws one 16-bit uniform per level from Philox(ctr = {e_lo, e_hi,
 * 'RMAT', j}, key = seed): call j supplies levels 8j..8j+7, low half-word of
 * out[w] first.  Level i (i = 0 is the most significant bit) picks quadrant
 * a (src 0, dst 0) if u < ta, b (0, 1) if u < tab, c (1, 0) if u < tabc,
 * else d (1, 1).  Thresholds are cumulative probabilities times 65536.
 * Self-loops and duplicate edges are kept; no vertex permutation. */
void gwo_rmat_edges(uint32_t scale, uint64_t e_begin, uint64_t e_end, uint64_t seed,
                    uint32_t ta, uint32_t tab, uint32_t tabc,
                    uint32_t *src, uint32_t *dst) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int64_t n = (int64_t)(e_end - e_begin);
    #pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        uint64_t e = e_begin + (uint64_t)k;
        uint32_t s = 0, d = 0, r[4];
        for (uint32_t lvl = 0; lvl < scale; ++lvl) {
            if ((lvl & 7u) == 0) {
                uint32_t c[4] = {(uint32_t)e, (uint32_t)(e >> 32), GWO_TAG_RMAT, lvl >> 3};
                philox(c, key, r);
            }
            uint32_t w = r[(lvl & 7u) >> 1];
            uint32_t u = (lvl & 1u) ? (w >> 16) : (w & 0xFFFFu);
            uint32_t sb = (u >= tab) ? 1u : 0u;             /* quadrants c, d */
            uint32_t db = (u >= ta && u < tab) || (u >= tabc) ? 1u : 0u; /* b, d */
            s = (s << 1) | sb;
            d = (d << 1) | db;
        }
        src[k] = s;
        dst[k] = d;
    }
}

/* Sources for multi-source workloads: source k = mulhi32(r0, n) with
 * r = Philox(ctr = {k, 0, 'SRCS', 0}, key = seed). */
void gwo_sources(uint64_t seed, uint32_t n, uint32_t count, uint32_t *out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t c[4] = {k, 0, GWO_TAG_SRCS, 0}, r[4];
        philox(c, key, r);
        out[k] = mulhi32(r[0], n);
    }
}

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* CSR from an edge list: off[n+1] (uint64, absolute), col[m] (uint32), each
 * row sorted ascending so the layout is independent of edge order.  Same
 * layout as conversions.hpp:181-298 writes to file_0.beg_pos / file_0.csr,
 * except that self-loops are kept. */
int gwo_csr_build(uint32_t n, uint64_t m, const uint32_t *src, const uint32_t *dst,
                  uint64_t *off, uint32_t *col) {
    memset(off, 0, sizeof(uint64_t) * ((size_t)n + 1));
    for (uint64_t e = 0; e < m; ++e) {
        if (src[e] >= n || dst[e] >= n) return -1;
        off[src[e] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    if (!cur) return -2;
    memcpy(cur, off, sizeof(uint64_t) * (size_t)n);
    for (uint64_t e = 0; e < m; ++e) col[cur[src[e]]++] = dst[e];
    free(cur);
    #pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        uint64_t b = off[v], e = off[v + 1];
        if (e - b > 1) qsort(col + b, (size_t)(e - b), sizeof(uint32_t), cmp_u32);
    }
    return 0;
}

/* ------------------------------------------------------------- stepping */
typedef struct {
    const uint64_t *off;
    const uint32_t *col;
    uint32_t n;
} gwo_graph;

typedef struct {
    uint32_t *vertex_counts; /* nullable, u32[n], incremented atomically per visit */
    uint64_t steps;          /* hop increments */
    uint64_t visits;         /* updateInfo calls */
    uint64_t hash;           /* FNV-1a over 64-bit words (vertex << 32 | hop) */
    /* optional per-walk trace */
    uint32_t *tr_v, *tr_hop;
    uint64_t tr_cap, tr_len;
} gwo_rec;

#define GWO_HASH_INIT 0xCBF29CE484222325ull
#define GWO_HASH_PRIME 0x100000001B3ull

static inline void record(gwo_rec *rc, uint32_t v, uint32_t hop) {
    rc->visits++;
    rc->hash = (rc->hash ^ (((uint64_t)v << 32) | hop)) * GWO_HASH_PRIME;
    if (rc->vertex_counts) __atomic_fetch_add(&rc->vertex_counts[v], 1u, __ATOMIC_RELAXED);
    if (rc->tr_v && rc->tr_len < rc->tr_cap) {
        rc->tr_v[rc->tr_len] = v;
        rc->tr_hop[rc->tr_len] = hop;
    }
    rc->tr_len++;
}

/* Advance one walk while its vertex lies in [lo, hi).  Returns GWO_FINISHED
 * when the walk ends (cap reached, died, or joint period complete), otherwise
 * GWO_MOVED with w->vertex outside [lo, hi).  Mirrors the while-loops of the
 * updateByWalk variants; the else-branch per kind:
 *   JUMP    teleport to mulhi32(r3, N)     (randomwalkwithjump.hpp:49-51)
 *   STOP    die                              (randomwalkwithstop.hpp:39-41)
 *   PROB    die, no hop cap                  (randomwalkwithprob.hpp:30,39-42)
 *   JOINT   restart at the source vertex; finish when hop % L == L-1
 *                                            (randomwalkwithrestartwithjoint.hpp:47-53)
 *   RESTART restart at the source vertex, cap L (randomwalkwithrestart.hpp:33-63) */
static int advance(const gwo_program *p, const gwo_graph *g, gwo_walk *w,
                   uint32_t lo, uint32_t hi, gwo_rec *rc) {
    const uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    const int capped = (p->kind == GWO_JUMP || p->kind == GWO_STOP || p->kind == GWO_RESTART);
    uint32_t v = w->vertex, hop = w->hop;
    while (v >= lo && v < hi) {
        if (capped && hop >= p->L) { w->vertex = v; w->hop = hop; return GWO_FINISHED; }
        record(rc, v, hop);
        uint32_t c[4] = {w->id, hop, GWO_TAG_WALK, p->stream}, r[4];
        philox(c, key, r);
        uint64_t b = g->off[v], deg = g->off[v + 1] - b;
        if (deg > 0 && (uint64_t)r[0] >= p->stop_thr) {
            uint64_t x = (uint64_t)r[1] | ((uint64_t)r[2] << 32);
            v = g->col[b + mulhi64(x, deg)];
        } else {
            switch (p->kind) {
            case GWO_JUMP: v = mulhi32(r[3], p->nvertices); break;
            case GWO_STOP:
            case GWO_PROB: w->vertex = v; w->hop = hop; return GWO_FINISHED;
            default: v = w->source; break; /* RESTART_JOINT, RESTART */
            }
        }
        hop++;
        rc->steps++;
        if (p->kind == GWO_RESTART_JOINT && hop % p->L == p->L - 1) {
            w->vertex = v; w->hop = hop; return GWO_FINISHED;
        }
    }
    w->vertex = v;
    w->hop = hop;
    if (capped && hop >= p->L) return GWO_FINISHED;
    return GWO_MOVED;
}

/* Whole-graph replay (one block spanning all vertices): every walk runs to
 * completion.  Per-walk outputs are optional (nullable arrays of nwalks).
 * Totals go to totals[0] = steps, totals[1] = visits. */
int gwo_replay(const gwo_program *p, const uint64_t *off, const uint32_t *col, uint32_t n,
               const gwo_walk *walks, uint64_t nwalks,
               uint32_t *vertex_counts, gwo_walk *final_state, uint64_t *walk_hash,
               uint32_t *walk_visits, uint64_t *totals) {
    gwo_graph g = {off, col, n};
    uint64_t steps = 0, visits = 0;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:steps, visits)
    for (int64_t i = 0; i < (int64_t)nwalks; ++i) {
        gwo_walk w = walks[i];
        gwo_rec rc;
        memset(&rc, 0, sizeof(rc));
        rc.vertex_counts = vertex_counts;
        rc.hash = GWO_HASH_INIT;
        if (w.vertex >= n) { if (final_state) final_state[i] = w; continue; }
        advance(p, &g, &w, 0, n, &rc);
        steps += rc.steps;
        visits += rc.visits;
        if (final_state) final_state[i] = w;
        if (walk_hash) walk_hash[i] = rc.hash;
        if (walk_visits) walk_visits[i] = (uint32_t)rc.visits;
    }
    if (totals) { totals[0] = steps; totals[1] = visits; }
    return 0;
}

/* Trace one walk: visited (vertex, hop) pairs in order.  Returns the number
 * of visits (which may exceed cap; only cap are stored). */
uint64_t gwo_trace(const gwo_program *p, const uint64_t *off, const uint32_t *col, uint32_t n,
                   const gwo_walk *walk, uint32_t *tr_v, uint32_t *tr_hop, uint64_t cap,
                   gwo_walk *final_state) {
    gwo_graph g = {off, col, n};
    gwo_walk w = *walk;
    gwo_rec rc;
    memset(&rc, 0, sizeof(rc));
    rc.hash = GWO_HASH_INIT;
    rc.tr_v = tr_v; rc.tr_hop = tr_hop; rc.tr_cap = cap;
    advance(p, &g, &w, 0, n, &rc);
    if (final_state) *final_state = w;
    return rc.tr_len;
}

/* --------------------------------------------------------- block engine */
/* Block of v: index b with block_first[b] <= v < block_first[b+1], or
 * nblocks if v is out of range (randomwalk.hpp:74-82 returns nblocks and the
 * caller drops the walk). */
static inline uint32_t block_of(const uint32_t *bf, uint32_t nblocks, uint32_t v) {
    if (v >= bf[nblocks]) return nblocks;
    uint32_t lo = 0, hi = nblocks; /* invariant: bf[lo] <= v < bf[hi] */
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (bf[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

/* One block round: every walk of bucket `b` advances inside the block.
 * status[i] = GWO_MOVED (walk now in another block) or GWO_FINISHED.
 * Per-walk hash/visit accumulators are carried across rounds by the caller. */
int gwo_block_round(const gwo_program *p, const uint64_t *off, const uint32_t *col, uint32_t n,
                    const uint32_t *block_first, uint32_t nblocks, uint32_t b,
                    gwo_walk *walks, uint64_t nwalks, uint8_t *status,
                    uint32_t *vertex_counts, uint64_t *totals) {
    gwo_graph g = {off, col, n};
    uint32_t lo = block_first[b], hi = block_first[b + 1];
    uint64_t steps = 0, visits = 0;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:steps, visits)
    for (int64_t i = 0; i < (int64_t)nwalks; ++i) {
        gwo_rec rc;
        memset(&rc, 0, sizeof(rc));
        rc.vertex_counts = vertex_counts;
        rc.hash = GWO_HASH_INIT;
        int st = advance(p, &g, &walks[i], lo, hi, &rc);
        if (st == GWO_MOVED && block_of(block_first, nblocks, walks[i].vertex) >= nblocks) st = GWO_FINISHED;
        status[i] = (uint8_t)st;
        steps += rc.steps;
        visits += rc.visits;
    }
    if (totals) { totals[0] += steps; totals[1] += visits; }
    return 0;
}

/* Stable counting sort of walk records by the block of their vertex
 * (the bucket gather of walk.hpp:114-137 without its per-thread order).
 * counts[b] receives the bucket sizes (nblocks + 1 entries: the last counts
 * out-of-range walks, which are dropped from `out`). */
void gwo_bucket_stable(const gwo_walk *in, uint64_t n, const uint32_t *block_first,
                       uint32_t nblocks, gwo_walk *out, uint64_t *counts) {
    memset(counts, 0, sizeof(uint64_t) * ((size_t)nblocks + 1));
    uint32_t *bk = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint64_t i = 0; i < n; ++i) {
        bk[i] = block_of(block_first, nblocks, in[i].vertex);
        counts[bk[i]]++;
    }
    uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)nblocks + 1));
    uint64_t acc = 0;
    for (uint32_t b = 0; b <= nblocks; ++b) { pos[b] = acc; acc += counts[b]; }
    for (uint64_t i = 0; i < n; ++i)
        if (bk[i] < nblocks) out[pos[bk[i]]++] = in[i];
    free(pos);
    free(bk);
}

/* ------------------------------------------------ FP64 stop-walk visits */
/* Expected visits of `walks` STOP walks (follow a uniform CSR slot with
 * probability 1 - p where the vertex has edges, else die; cap L hops,
 * randomwalkwithstop.hpp:20-52) from each of K sources: out[v * K + k] =
 * sum_{h < L} x_h[v] for source k, x_0 = walks at the source and
 * x_{h+1}[d] = (1 - p) sum over the in-edges (u -> d) of x_h[u] / deg(u),
 * computed by pulling over the transposed CSR (toff, tcol: the in-neighbours
 * of each vertex, one entry per edge) with deg the forward out-degrees.
 * The same chain as fp64.expected_visits for a graph too large for a sparse
 * matrix in Python; K columns share each edge's read. */
void gwo_fp64_stop_visits(uint32_t n, const uint64_t *toff, const uint32_t *tcol, const uint64_t *off,
                          uint32_t K, const uint32_t *sources, double walks, uint32_t L, double p, double *out) {
    const size_t nk = (size_t)n * K;
    double *x = (double *)calloc(nk ? nk : 1, sizeof(double));
    double *y = (double *)calloc(nk ? nk : 1, sizeof(double));
    double *inv = (double *)malloc(sizeof(double) * (n ? n : 1));
    #pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
        const uint64_t d = off[v + 1] - off[v];
        inv[v] = d ? (1.0 - p) / (double)d : 0.0;
    }
    memset(out, 0, sizeof(double) * nk);
    for (uint32_t k = 0; k < K; ++k) x[(size_t)sources[k] * K + k] += walks;
    for (uint32_t h = 0; h < L; ++h) {
        #pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            double *o = out + (size_t)v * K;
            const double *xv = x + (size_t)v * K;
            for (uint32_t k = 0; k < K; ++k) o[k] += xv[k];
        }
        if (h + 1 == L) break;
        #pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t d = 0; d < (int64_t)n; ++d) {
            double acc[64];
            for (uint32_t k = 0; k < K; ++k) acc[k] = 0.0;
            for (uint64_t e = toff[d]; e < toff[d + 1]; ++e) {
                const uint32_t u = tcol[e];
                const double w = inv[u];
                const double *xu = x + (size_t)u * K;
                for (uint32_t k = 0; k < K; ++k) acc[k] += xu[k] * w;
            }
            double *yd = y + (size_t)d * K;
            for (uint32_t k = 0; k < K; ++k) yd[k] = acc[k];
        }
        double *t = x;
        x = y;
        y = t;
    }
    free(x);
    free(y);
    free(inv);
}

int gwo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void gwo_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
