// engine/graphwalker.hpp -- graphwalker_engine: the block loop the apps drive
// (src/engine/graphwalker.hpp:26-311), running on one MI355X through the C ABI
// of libgwhip (include/gwhip.h).
//
//   graphwalker_engine engine(file, blocksize_kb, nblocks, nmblocks, m);
//   engine.run(program, prob);
//
// The constructor loads the preprocessed CSR and block table.  In HBM by
// default (the whole graph: the device block cache is the HBM); with option
// outofcore 1 -- or by default when the CSR would not fit three quarters of
// the device memory -- the CSR stays in pinned host memory and runs step it
// through a device block cache of nmblocks slots, the reference's memory
// budget (graphwalker.hpp:84-100, findSubGraph / swapOut :168-251).
// run() calls the program's startWalks, gathers the per-(thread, block)
// buckets into device walk records, runs the program's step semantics on the
// GPU with a per-walk event log, and replays every visit to the program's
// updateInfo on the host in walk order -- the same calls the reference makes
// per step, so app recorders see the same (source, vertex, hop) visits.
//
// Options (get_option_*): execthreads, schedule (persistent|blocked|plain),
// seed (Philox key), device, outofcore (-1 auto, 0 in HBM, 1 host-resident),
// and the recorder the app's updateInfo is fed from:
//   recorder events  (default) the per-walk event log: every visit replayed
//                    with its (source, vertex, hop), walk by walk;
//   recorder source  per-(source, vertex) counts on the device
//                    (GW_REC_SOURCE_COUNT), replayed as that many
//                    updateInfo(source, vertex, t, 0) calls -- for apps whose
//                    updateInfo depends on the source and the vertex only
//                    (msppr's visit distributions);
//   recorder vertex  per-vertex counts on the device (GW_REC_VERTEX_COUNT),
//                    replayed as updateInfo(vertex, vertex, t, 0) calls -- for
//                    apps that count visits per vertex (rwdomination);
//   recorder last    each walk's last visit only (GW_REC_EVENTS_LAST with a
//                    one-visit log), replayed with its (source, vertex, hop) --
//                    for apps whose updateInfo returns early below the final
//                    hop (graphlet counts walks back at their source at hop
//                    L-1);
//   recorder none    no replay (an app whose updateInfo does nothing, such as
//                    rawrandomwalks).
// replaythreads (default 1) replays on that many threads, each with its own
// threadid (< execthreads) and a contiguous range of sources (or vertices),
// so per-thread recorders see disjoint sources; 1 replays in walk order.
#ifndef DEF_GRAPHCHWALKER_ENGINE
#define DEF_GRAPHCHWALKER_ENGINE

#include <assert.h>
#include <fcntl.h>
#include <omp.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <fstream>
#include <parallel/algorithm>
#include <string>
#include <vector>

#include "api/cmdopts.hpp"
#include "api/datatype.hpp"
#include "api/filename.hpp"
#include "api/io.hpp"
#include "gwhip.h"
#include "logger/logger.hpp"
#include "metrics/metrics.hpp"
#include "walks/randomwalk.hpp"

class graphwalker_engine {
public:
    std::string base_filename;
    unsigned long long blocksize_kb;
    bid_t nblocks;
    vid_t nvertices;
    tid_t exec_threads;
    vid_t* blocks;
    timeval start;
    bid_t nmblocks;
    bid_t exec_block;
    metrics& m;
    WalkManager* walk_manager;

    graphwalker_engine(std::string _base_filename, unsigned long long _blocksize_kb, bid_t _nblocks, bid_t _nmblocks,
                       metrics& _m)
        : base_filename(_base_filename), blocksize_kb(_blocksize_kb), nblocks(_nblocks), nvertices(0), blocks(nullptr),
          nmblocks(_nmblocks), exec_block(0), m(_m), walk_manager(nullptr), ctx_(nullptr) {
        exec_threads = static_cast<tid_t>(get_option_int("execthreads", omp_get_max_threads()));
        omp_set_num_threads(exec_threads);
        load_block_range();
        load_graph();
        walk_manager = new WalkManager(m, nblocks, exec_threads, base_filename);
        m.set("file", base_filename);
        m.set("engine", std::string("gwhip"));
        m.set("nblocks", static_cast<size_t>(nblocks));
        logstream(LOG_INFO) << "Engine configuration: exec_threads = " << static_cast<int>(exec_threads)
                            << ", blocksize_kb = " << blocksize_kb << ", nblocks = " << nblocks
                            << ", nvertices = " << nvertices << std::endl;
    }

    virtual ~graphwalker_engine() {
        delete walk_manager;
        free(blocks);
        if (ctx_) gw_destroy(ctx_);
    }

    graphwalker_engine(const graphwalker_engine&) = delete;
    graphwalker_engine& operator=(const graphwalker_engine&) = delete;

    virtual size_t num_vertices() { return blocks[nblocks]; }

    gw_ctx* context() { return ctx_; }

    // `prob` chooses between the min-step and max-walks block heuristics
    // (walk.hpp:245-252).  With the graph resident in HBM every block is
    // stepped each round, so it does not apply; out of core it picks the
    // blocks each load pass brings into the device block cache.
    void run(RandomWalk& userprogram, float prob) {
        if (host_resident_)
            check(gw_stream_policy(ctx_, GW_BLOCKS_MIN_STEP, prob, static_cast<uint64_t>(get_option_long("seed", 1))),
                  "gw_stream_policy");
        m.start_time("0_startWalks");
        userprogram.startWalks(*walk_manager, nblocks, blocks, base_filename);
        m.stop_time("0_startWalks");
        gettimeofday(&start, nullptr);
        m.start_time("00_runtime");

        m.start_time("3_getCurrentWalks");
        std::vector<gw_walk> walks = gather_walks();
        m.stop_time("3_getCurrentWalks");

        const gw_device_walk dp = userprogram.device_program();
        gw_program prog;
        memset(&prog, 0, sizeof prog);
        prog.kind = dp.kind;
        prog.L = dp.L;
        prog.stop_thr = static_cast<uint64_t>(llround(static_cast<double>(dp.p) * 4294967296.0));
        prog.nvertices = dp.kind == GW_JUMP ? (dp.N ? dp.N : nvertices) : 0;
        prog.seed = get_option_long("seed", 1);
        const std::string recm = get_option_string("recorder", "events");
        if (recm != "events" && recm != "last" && recm != "source" && recm != "vertex" && recm != "none")
            logstream(LOG_FATAL) << "recorder must be events, last, source, vertex or none (got " << recm << ")"
                                 << std::endl;
        const bool last_only = recm == "last";
        prog.recorder = recm == "events" ? (GW_REC_EVENTS | GW_REC_WALK_SUMMARY)
                      : last_only        ? (GW_REC_EVENTS | GW_REC_WALK_SUMMARY | GW_REC_EVENTS_LAST)
                      : recm == "source" ? (GW_REC_SOURCE_COUNT | GW_REC_NO_FINAL)
                      : recm == "vertex" ? (GW_REC_VERTEX_COUNT | GW_REC_NO_FINAL)
                                         : GW_REC_NO_FINAL;
        const std::string sched = get_option_string("schedule", "persistent");
        prog.schedule = sched == "plain" ? GW_SCHED_PLAIN : sched == "blocked" ? GW_SCHED_BLOCKED : GW_SCHED_PERSISTENT;
        if (prog.schedule == GW_SCHED_BLOCKED && !blocks_on_device_) prog.schedule = GW_SCHED_PERSISTENT;
        if (host_resident_) prog.schedule = GW_SCHED_BLOCKED;  // the block cache steps block by block
        prog.event_cap = last_only ? 1u : dp.kind == GW_PROB ? 64u : std::max<uint32_t>(1u, dp.L);

        m.start_time("2_findSubGraph");
        check(gw_walks_load(ctx_, walks.data(), walks.size()), "gw_walks_load");
        m.stop_time("2_findSubGraph");

        gw_stats st;
        const int rthreads = std::max(1, std::min<int>(get_option_int("replaythreads", 1), exec_threads));
        if (recm == "events" || last_only) {
            std::vector<uint64_t> events;
            std::vector<uint32_t> visits(walks.size());
            for (;;) {
                m.start_time("5_exec_updates");
                check(gw_run(ctx_, &prog, &st), "gw_run");
                m.stop_time("5_exec_updates");
                check(gw_summary_read(ctx_, nullptr, visits.data(), visits.size()), "gw_summary_read");
                const uint32_t vmax = visits.empty() ? 0 : *std::max_element(visits.begin(), visits.end());
                if (last_only || vmax <= prog.event_cap) break;  // the last-visit log never outgrows its slot
                // an uncapped walk outgrew the log: rerun (the same walks) with a
                // log that holds the longest, if the device holds it
                uint64_t hbm = 0;
                check(gw_device_info(ctx_, nullptr, 0, nullptr, &hbm), "gw_device_info");
                const uint64_t need = static_cast<uint64_t>(walks.size()) * vmax * 8;
                if (need > hbm / 2)
                    logstream(LOG_FATAL) << "the event log for walks of up to " << vmax << " visits needs " << need
                                         << " bytes, over half the device memory; run with recorder source, vertex "
                                            "or none" << std::endl;
                prog.event_cap = vmax;
            }
            events.resize(walks.size() * static_cast<size_t>(prog.event_cap));
            if (!walks.empty())
                check(gw_events_read(ctx_, events.data(), walks.size(), prog.event_cap), "gw_events_read");
            // Replay every logged visit (with recorder last, each walk's last
            // one) to updateInfo, walk by walk, hop by hop; with
            // replaythreads > 1 each thread takes a contiguous range of the
            // walks (sorted by source) cut at source boundaries.
            m.start_time("z_replay_updateInfo");
            const std::vector<size_t> cut = split_by_key(walks.size(), rthreads, [&](size_t i) { return walks[i].source; });
#pragma omp parallel for num_threads(rthreads) schedule(static, 1)
            for (int t = 0; t < rthreads; ++t) {
                for (size_t i = cut[t]; i < cut[t + 1]; ++i) {
                    const gw_walk& w = walks[i];
                    const uint64_t* ev = events.data() + static_cast<size_t>(w.id) * prog.event_cap;
                    const uint32_t nk = last_only ? std::min<uint32_t>(visits[w.id], 1u) : visits[w.id];
                    for (uint32_t k = 0; k < nk; ++k)
                        userprogram.updateInfo(w.source, static_cast<vid_t>(ev[k] & 0xffffffffu), static_cast<tid_t>(t),
                                               static_cast<hid_t>(ev[k] >> 32));
                }
            }
            m.stop_time("z_replay_updateInfo");
        } else {
            m.start_time("5_exec_updates");
            check(gw_run(ctx_, &prog, &st), "gw_run");
            m.stop_time("5_exec_updates");
            m.start_time("z_replay_updateInfo");
            if (recm == "vertex") {
                std::vector<uint32_t> counts(nvertices);
                check(gw_counts_read(ctx_, counts.data(), nvertices), "gw_counts_read");
                const std::vector<size_t> cut = split_by_key(nvertices, rthreads, [](size_t i) { return i; });
#pragma omp parallel for num_threads(rthreads) schedule(static, 1)
                for (int t = 0; t < rthreads; ++t)
                    for (size_t v = cut[t]; v < cut[t + 1]; ++v)
                        for (uint32_t c = 0; c < counts[v]; ++c)
                            userprogram.updateInfo(static_cast<vid_t>(v), static_cast<vid_t>(v), static_cast<tid_t>(t), 0);
            } else if (recm == "source") {
                uint64_t np = 0;
                check(gw_source_counts_read(ctx_, nullptr, nullptr, nullptr, 0, &np), "gw_source_counts_read");
                std::vector<uint32_t> ps(np), pv(np), pc(np);
                check(gw_source_counts_read(ctx_, ps.data(), pv.data(), pc.data(), np, &np), "gw_source_counts_read");
                const std::vector<size_t> cut = split_by_key(np, rthreads, [&](size_t i) { return ps[i]; });
#pragma omp parallel for num_threads(rthreads) schedule(static, 1)
                for (int t = 0; t < rthreads; ++t)
                    for (size_t i = cut[t]; i < cut[t + 1]; ++i)
                        for (uint32_t c = 0; c < pc[i]; ++c)
                            userprogram.updateInfo(ps[i], pv[i], static_cast<tid_t>(t), 0);
            }
            m.stop_time("z_replay_updateInfo");
        }
        m.set("walk_steps", static_cast<size_t>(st.steps));
        m.set("walk_visits_device", static_cast<size_t>(st.visits));
        m.set("device_ms", st.total_ms);
        if (host_resident_) {
            m.set("blocks_loaded", static_cast<size_t>(st.blocks_loaded));
            m.set("h2d_bytes", static_cast<size_t>(st.h2d_bytes));
        }

        walk_manager->walksum = 0;
        for (bid_t b = 0; b < nblocks; ++b) {
            walk_manager->walknum[b] = 0;
            walk_manager->minstep[b] = 0xffff;
        }
        m.stop_time("00_runtime");
    }

private:
    gw_ctx* ctx_;

    // Cut [0, n) into `parts` contiguous ranges of about equal size whose
    // boundaries fall where key(i) changes (key non-decreasing in i), so no
    // key spans two ranges.  Returns parts + 1 boundaries.
    template <class Key>
    static std::vector<size_t> split_by_key(size_t n, int parts, Key key) {
        std::vector<size_t> cut(parts + 1, n);
        cut[0] = 0;
        for (int t = 1; t < parts; ++t) {
            size_t i = std::max(cut[t - 1], n * t / parts);
            while (i > cut[t - 1] && i < n && key(i) == key(i - 1)) ++i;
            cut[t] = i;
        }
        return cut;
    }
    bool blocks_on_device_ = false;
    bool host_resident_ = false;

    void check(int rc, const char* what) {
        if (rc != GW_OK) logstream(LOG_FATAL) << what << " failed (" << rc << "): " << gw_last_error(ctx_) << std::endl;
    }

    void load_block_range() {
        const std::string f = blockrangename(base_filename, blocksize_kb);
        std::ifstream in(f.c_str());
        if (!in.good()) logstream(LOG_FATAL) << "Could not load block range file: " << f << std::endl;
        blocks = static_cast<vid_t*>(malloc((nblocks + 1) * sizeof(vid_t)));
        for (bid_t i = 0; i <= nblocks; ++i) {
            if (!(in >> blocks[i])) logstream(LOG_FATAL) << "block range file too short: " << f << std::endl;
        }
    }

    void load_graph() {
        const std::string base = fidname(base_filename, 0);
        gw_cfg cfg;
        memset(&cfg, 0, sizeof cfg);
        cfg.device = get_option_int("device", 0);
        if (gw_create(&ctx_, &cfg) != GW_OK) logstream(LOG_FATAL) << "no usable HIP device" << std::endl;
        const size_t off_bytes = file_bytes(base + ".beg_pos"), col_bytes = file_bytes(base + ".csr");
        if (off_bytes < 2 * sizeof(eid_t)) logstream(LOG_FATAL) << "empty beg_pos file " << base << ".beg_pos" << std::endl;
        nvertices = static_cast<vid_t>(off_bytes / sizeof(eid_t) - 1);
        const uint64_t m_edges = col_bytes / sizeof(vid_t);
        int ooc = get_option_int("outofcore", -1);
        if (ooc < 0) {
            uint64_t hbm = 0;
            check(gw_device_info(ctx_, nullptr, 0, nullptr, &hbm), "gw_device_info");
            ooc = off_bytes + col_bytes > hbm / 4 * 3 ? 1 : 0;
        }
        // The block table of the reference cut rule may hold empty blocks;
        // the device block loop takes non-empty ones only.
        std::vector<uint32_t> bf;
        for (bid_t b = 0; b <= nblocks; ++b)
            if (bf.empty() || blocks[b] > bf.back()) bf.push_back(blocks[b]);
        m.start_time("g_loadSubGraph");
        if (ooc) {
            // read straight into pinned memory the engine uses in place
            eid_t* off = nullptr;
            vid_t* col = nullptr;
            check(gw_host_alloc(ctx_, off_bytes, reinterpret_cast<void**>(&off)), "gw_host_alloc");
            check(gw_host_alloc(ctx_, std::max<size_t>(col_bytes, 4), reinterpret_cast<void**>(&col)), "gw_host_alloc");
            read_into(base + ".beg_pos", off, off_bytes);
            read_into(base + ".csr", col, col_bytes);
            // nmblocks slots of the largest block (offsets 256-byte aligned, then columns)
            uint64_t slot = 0;
            for (size_t b = 0; b + 1 < bf.size(); ++b) {
                const uint64_t ob = ((bf[b + 1] - bf[b] + 1ull) * 8 + 255) / 256 * 256;
                slot = std::max<uint64_t>(slot, (ob + (off[bf[b + 1]] - off[bf[b]]) * 4 + 255) / 256 * 256);
            }
            const uint64_t cache = std::max<uint64_t>(1, nmblocks) * std::max<uint64_t>(slot, 256);
            check(gw_graph_set_host(ctx_, off, col, nvertices, m_edges, cache), "gw_graph_set_host");
            host_resident_ = true;
            if (bf.back() == nvertices && bf.size() >= 2 && bf.size() - 1 <= 2048)
                check(gw_graph_blocks(ctx_, bf.data(), static_cast<uint32_t>(bf.size() - 1)), "gw_graph_blocks");
            blocks_on_device_ = true;
        } else {
            std::vector<eid_t> off;
            std::vector<vid_t> col;
            read_file(base + ".beg_pos", off);
            read_file(base + ".csr", col);
            check(gw_graph_set(ctx_, off.data(), col.data(), nvertices, col.size()), "gw_graph_set");
            if (bf.back() == nvertices && bf.size() >= 2 && bf.size() - 1 <= 2048)
                blocks_on_device_ = gw_graph_blocks(ctx_, bf.data(), static_cast<uint32_t>(bf.size() - 1)) == GW_OK;
        }
        m.stop_time("g_loadSubGraph");
    }

    static size_t file_bytes(const std::string& f) {
        int fd = open(f.c_str(), O_RDONLY);
        if (fd < 0) logstream(LOG_FATAL) << "Could not load :" << f << ", error: " << strerror(errno) << std::endl;
        const off_t bytes = lseek(fd, 0, SEEK_END);
        close(fd);
        return static_cast<size_t>(bytes);
    }

    static void read_into(const std::string& f, void* out, size_t bytes) {
        int fd = open(f.c_str(), O_RDONLY);
        if (fd < 0) logstream(LOG_FATAL) << "Could not load :" << f << ", error: " << strerror(errno) << std::endl;
        if (bytes) preada(fd, static_cast<char*>(out), bytes, 0);
        close(fd);
    }

    template <class T>
    static void read_file(const std::string& f, std::vector<T>& out) {
        int fd = open(f.c_str(), O_RDONLY);
        if (fd < 0) logstream(LOG_FATAL) << "Could not load :" << f << ", error: " << strerror(errno) << std::endl;
        const off_t bytes = lseek(fd, 0, SEEK_END);
        out.resize(static_cast<size_t>(bytes) / sizeof(T));
        if (!out.empty()) preada(fd, out.data(), out.size() * sizeof(T), 0);
        close(fd);
    }

    // Buckets -> device records.  Walk ids follow the order of the packed
    // records sorted by value (then block), so they do not depend on which
    // thread placed which walk.
    std::vector<gw_walk> gather_walks() {
        struct rec { WalkDataType w; bid_t p; };
        size_t total = 0;
        for (bid_t p = 0; p < nblocks; ++p)
            for (tid_t t = 0; t < exec_threads; ++t) total += walk_manager->pwalks[t][p].size_w;
        std::vector<rec> all;
        all.reserve(total);
        for (bid_t p = 0; p < nblocks; ++p)
            for (tid_t t = 0; t < exec_threads; ++t) {
                WalkBuffer& b = walk_manager->pwalks[t][p];
                for (wid_t k = 0; k < b.size_w; ++k) all.push_back({b[k], p});
                b.size_w = 0;
            }
        // (record, block) pairs are unique up to identical walks, so any sort
        // by that pair gives the same ids: sorted in parallel on the host's cores
        __gnu_parallel::sort(all.begin(), all.end(), [](const rec& a, const rec& b) {
            return a.w != b.w ? a.w < b.w : a.p < b.p;
        });
        std::vector<gw_walk> out(all.size());
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < all.size(); ++i) {
            const WalkDataType w = all[i].w;
            out[i].id = static_cast<uint32_t>(i);
            out[i].source = walk_manager->getSourceId(w);
            out[i].vertex = blocks[all[i].p] + walk_manager->getCurrentId(w);
            out[i].hop = walk_manager->getHop(w);
        }
        return out;
    }
};

#endif
