// TEST INFRASTRUCTURE / CPU BASELINE: an app on the reference GraphWalker API
// (compiled by oracle/build_ref.sh against the reference's own headers into
// oracle/_ref/), so that the reference CPU engine can be timed on the batched
// multi-source PPR workload with uniformly drawn sources.  The reference's
// msppr app (apps/msppr.cpp:41-76) places walks on one contiguous range of
// source ids; this app places walkspersource walks on each vertex of a source
// list read from a file, and records visits as msppr does: one
// DiscreteDistribution per source (walks/discretedistribution.hpp:44-57)
// plus a per-thread used-edge count (msppr.cpp:71-76).  The distributions
// are kept per engine thread (msppr shares one per source between threads,
// which races on its size field); each visit does the same linear search.
//
// Walk program: the reference's RandomWalkwithStop (walks/randomwalkwithstop.hpp),
// stop probability 0.15 per hop, cap L.
// Options (command line, the reference's option parser):
//   file <base>  sources <path: u32 list>  walkspersource <W>  L <cap>
//   blocksize_kb <kb>  nmblocks <m>  execthreads <t>  out <path>
// Output: <path> holds, for source 0, the entries of its distribution summed
// over threads as u32 (vertex, count) pairs; the engine's own metrics file
// holds the timers (00_runtime).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "api/graphwalker_basic_includes.hpp"
#include "walks/discretedistribution.hpp"
#include "walks/randomwalkwithstop.hpp"

class SourceListPPR : public RandomWalkwithStop {
public:
    std::vector<vid_t> sources;
    wid_t walkspersource = 0;
    tid_t nthreads = 1;
    std::vector<DiscreteDistribution*> freq;  // [thread][source]
    std::vector<eid_t> used_edges;

    void setup(const std::vector<vid_t>& src, wid_t wps, hid_t cap) {
        sources = src;
        walkspersource = wps;
        initializeRW(static_cast<wid_t>(src.size()) * wps, cap);
    }

    void startWalksbyApp(WalkManager& walk_manager) {
        // the engine has set its thread count by now (graphwalker.hpp:76-77)
        nthreads = omp_get_max_threads();
        for (tid_t t = 0; t < nthreads; t++) freq.push_back(new DiscreteDistribution[sources.size()]);
        used_edges.assign(nthreads, 0);
        for (bid_t p = 0; p < nblocks; p++) {
            walk_manager.minstep[p] = 0;
            walk_manager.walknum[p] = 0;
        }
        // bucket the sources by block, then place each block's walks
        std::vector<std::vector<vid_t>> by_block(nblocks);
        for (vid_t k = 0; k < sources.size(); k++) by_block[getblock(sources[k])].push_back(k);
        walk_manager.walksum = 0;
        for (bid_t p = 0; p < nblocks; p++) {
            const std::vector<vid_t>& ks = by_block[p];
            walk_manager.walknum[p] = static_cast<wid_t>(ks.size()) * walkspersource;
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < ks.size(); i++) {
                const vid_t k = ks[i];
                const vid_t cur = sources[k] - blocks[p];
                const WalkDataType walk = walk_manager.encode(k, cur, 0);
                for (wid_t j = 0; j < walkspersource; j++) walk_manager.moveWalk(walk, p, omp_get_thread_num(), cur);
            }
            walk_manager.walksum += walk_manager.walknum[p];
        }
    }

    void updateInfo(vid_t s, vid_t dstId, tid_t threadid, hid_t hop) {
        freq[threadid][s].add(dstId);
        used_edges[threadid]++;
    }
};

int main(int argc, const char** argv) {
    set_argc(argc, argv);
    metrics m("msppr_sources");
    std::string base = get_option_string("file");
    std::string src_path = get_option_string("sources");
    std::string out = get_option_string("out", "");
    wid_t wps = get_option_long("walkspersource", 2000);
    hid_t cap = get_option_int("L", 20);
    unsigned long long blocksize_kb = get_option_long("blocksize_kb", 0);
    bid_t nmblocks = get_option_int("nmblocks", 0);
    float prob = get_option_float("prob", 0.2);

    std::vector<vid_t> sources;
    FILE* f = fopen(src_path.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "msppr_sources: cannot read %s\n", src_path.c_str());
        return 1;
    }
    uint32_t v;
    while (fread(&v, sizeof v, 1, f) == 1) sources.push_back(v);
    fclose(f);

    SourceListPPR program;
    program.setup(sources, wps, cap);
    if (blocksize_kb == 0) blocksize_kb = program.compBlockSize(static_cast<wid_t>(sources.size()) * wps);
    bid_t nblocks = convert_if_notexists(base, blocksize_kb);
    if (nmblocks == 0) nmblocks = program.compNmblocks(blocksize_kb);
    if (nmblocks > nblocks) nmblocks = nblocks;
    graphwalker_engine engine(base, blocksize_kb, nblocks, nmblocks, m);
    engine.run(program, prob);

    if (!out.empty()) {
        std::map<vid_t, uint64_t> c;
        for (tid_t t = 0; t < program.nthreads; t++) {
            DiscreteDistribution& d = program.freq[t][0];
            for (unsigned i = 0; i < d.size; i++) c[d.ids[i]] += d.counts[i];
        }
        FILE* o = fopen(out.c_str(), "wb");
        for (const auto& kv : c) {
            const uint32_t pair[2] = {kv.first, static_cast<uint32_t>(kv.second)};
            fwrite(pair, sizeof pair, 1, o);
        }
        fclose(o);
    }
    eid_t used = 0;
    for (eid_t u : program.used_edges) used += u;
    printf("visits: %llu\n", static_cast<unsigned long long>(used));
    metrics_report(m);
    return 0;
}
