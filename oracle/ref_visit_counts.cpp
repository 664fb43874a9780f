// TEST INFRASTRUCTURE: an app on the reference GraphWalker API (compiled by
// oracle/build_ref.sh against the reference's own headers into oracle/_ref/),
// so that the reference CPU engine's single-source PPR visit vector can be
// compared with the FP64 expectation.  The reference's msppr app
// (apps/msppr.cpp:73,138) keeps the same counts but prints only a top-20;
// this app writes the whole vector.
//
// Walk program: the reference's RandomWalkwithStop (walks/randomwalkwithstop.hpp),
// i.e. PPR by walks that end with probability 0.15 per hop or at cap L.
// Options (command line, the reference's option parser):
//   file <base>  N <vertices>  source <v>  walks <W>  L <cap>  out <path>
//   blocksize_kb <kb>  nmblocks <m>  execthreads <t>
// Output: <path> holds n little-endian u32 visit counts (updateInfo calls
// per vertex, randomwalk.hpp:32), summed over the engine's threads.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "api/graphwalker_basic_includes.hpp"
#include "walks/randomwalkwithstop.hpp"

class VisitCounts : public RandomWalkwithStop {
public:
    vid_t source = 0, nverts = 0;
    wid_t nwalks = 0;
    std::vector<std::vector<uint32_t>> per_thread;

    void setup(vid_t src, vid_t n, wid_t w, hid_t cap) {
        source = src;
        nverts = n;
        nwalks = w;
        initializeRW(w, cap);
    }

    void startWalksbyApp(WalkManager &walk_manager) {
        // the engine has set its thread count by now (graphwalker.hpp:76-77)
        per_thread.assign(omp_get_max_threads(), std::vector<uint32_t>(nverts, 0));
        bid_t p = getblock(source);
        vid_t local = source - blocks[p];
        walk_manager.minstep[p] = 0;
        walk_manager.walknum[p] = nwalks;
        const WalkDataType start = walk_manager.encode(0, local, 0);
#pragma omp parallel for schedule(static)
        for (wid_t j = 0; j < nwalks; j++) walk_manager.moveWalk(start, p, omp_get_thread_num(), local);
        walk_manager.walksum = nwalks;
    }

    void updateInfo(vid_t s, vid_t v, tid_t t, hid_t hop) { per_thread[t][v]++; }

    std::vector<uint32_t> total() const {
        std::vector<uint32_t> c(nverts, 0);
        for (const auto &t : per_thread)
            for (vid_t v = 0; v < nverts; v++) c[v] += t[v];
        return c;
    }
};

int main(int argc, const char **argv) {
    set_argc(argc, argv);
    metrics m("visit_counts");
    std::string base = get_option_string("file");
    std::string out = get_option_string("out");
    vid_t source = get_option_int("source", 0);
    vid_t n = get_option_int("N");
    wid_t walks = get_option_long("walks", 100000);
    hid_t cap = get_option_int("L", 10);
    unsigned long long blocksize_kb = get_option_long("blocksize_kb", 0);
    bid_t nmblocks = get_option_int("nmblocks", 0);

    VisitCounts program;
    program.setup(source, n, walks, cap);
    if (blocksize_kb == 0) blocksize_kb = program.compBlockSize(walks);
    bid_t nblocks = convert_if_notexists(base, blocksize_kb);
    if (nmblocks == 0) nmblocks = program.compNmblocks(blocksize_kb);
    if (nmblocks > nblocks) nmblocks = nblocks;
    graphwalker_engine engine(base, blocksize_kb, nblocks, nmblocks, m);
    engine.run(program, 0.2f);

    std::vector<uint32_t> c = program.total();
    FILE *f = fopen(out.c_str(), "wb");
    if (!f || fwrite(c.data(), sizeof(uint32_t), c.size(), f) != c.size()) {
        fprintf(stderr, "visit_counts: cannot write %s\n", out.c_str());
        return 1;
    }
    fclose(f);
    metrics_report(m);
    return 0;
}
