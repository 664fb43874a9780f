// walkapp.cpp -- a walk app on the GraphWalker API (same shape as the
// reference apps: a RandomWalk subclass with startWalksbyApp + updateInfo,
// graphwalker_engine, metrics), used to check the C++ front end end to end.
//
//   walkapp file <g> program <jump|stop|prob|joint|restart> R <walks> L <hops>
//           [source <v>] [samesource 0|1] [walkprob 0.15] [seed 1]
//           [blocksize_kb <KB>] [out <counts file>]
//
// Walk i starts at vertex (source + i) mod N (all at `source` with
// samesource 1) with hop 0; its source field is its start vertex.  Every
// visit increments a per-vertex u32 count (rwdomination.cpp:63-66); the
// counts are written to `out` as raw uint32.
#include <string>
#include <type_traits>
#include <vector>

#include "api/graphwalker_basic_includes.hpp"
#include "walks/randomwalkwithjump.hpp"
#include "walks/randomwalkwithprob.hpp"
#include "walks/randomwalkwithrestart.hpp"
#include "walks/randomwalkwithrestartwithjoint.hpp"
#include "walks/randomwalkwithstop.hpp"
#include "util/toplist.hpp"

template <class Base>
class CountingWalks : public Base {
public:
    vid_t N = 0, first = 0;
    bool same = false;
    wid_t nwalks = 0;
    std::vector<VertexDataType> counts;
    wid_t visits = 0;

    void setup(vid_t _N, wid_t _R, hid_t _L, vid_t _first, bool _same, float p) {
        N = _N;
        nwalks = _R;
        first = _first;
        same = _same;
        if constexpr (std::is_same<Base, RandomWalkwithJump>::value) this->initializeRW(_N, _R, _L);
        else this->initializeRW(_R, _L);
        this->walkprob = p;
        counts.assign(N, 0);
    }

    void startWalksbyApp(WalkManager& wm) override {
        const tid_t nthreads = wm.num_threads();
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
        for (wid_t i = 0; i < nwalks; ++i) {
            const vid_t s = same ? first : static_cast<vid_t>((first + i) % N);
            const bid_t p = this->getblock(s);
            const vid_t cur = s - this->blocks[p];
            const WalkDataType w = wm.encode(s, cur, 0);
            wm.moveWalk(w, p, static_cast<tid_t>(omp_get_thread_num()), cur);
        }
        wm.walksum = nwalks;
    }

    void updateInfo(vid_t s, vid_t dst, tid_t tid, hid_t hop) override {
        (void)s; (void)tid; (void)hop;
        counts[dst]++;
        ++visits;
    }
};

template <class Prog>
int run_app(const std::string& file, vid_t N, wid_t R, hid_t L, vid_t source, bool same, float p, metrics& m,
            Prog& program) {
    program.setup(N, R, L, source, same, p);
    unsigned long long blocksize_kb = get_option_long("blocksize_kb", 0);
    if (blocksize_kb == 0) blocksize_kb = program.compBlockSize(R);
    const bid_t nblocks = convert_if_notexists(file, blocksize_kb);
    const bid_t nmblocks = static_cast<bid_t>(get_option_int("nmblocks", 1));
    graphwalker_engine engine(file, blocksize_kb, nblocks, nmblocks, m);
    engine.run(program, 0.2f);
    const std::string out = get_option_string("out", "");
    if (!out.empty()) {
        FILE* f = fopen(out.c_str(), "wb");
        if (!f || fwrite(program.counts.data(), sizeof(VertexDataType), N, f) != N) {
            logstream(LOG_ERROR) << "cannot write " << out << std::endl;
            return 1;
        }
        fclose(f);
    }
    std::cout << "walk_visits " << program.visits << std::endl;
    return 0;
}

int main(int argc, const char** argv) {
    set_argc(argc, argv);
    metrics m("walkapp");
    const std::string file = get_option_string("file");
    const std::string prog = get_option_string("program", "restart");
    const wid_t R = get_option_long("R", 1000);
    const hid_t L = static_cast<hid_t>(get_option_int("L", 10));
    const vid_t source = static_cast<vid_t>(get_option_int("source", 0));
    const bool same = get_option_int("samesource", 0) != 0;
    const float p = get_option_float("walkprob", 0.15f);
    convert_if_notexists(file, 1024);
    const vid_t N = static_cast<vid_t>(get_num_vertices(file));
    if (prog == "jump") { CountingWalks<RandomWalkwithJump> a; return run_app(file, N, R, L, source, same, p, m, a); }
    if (prog == "stop") { CountingWalks<RandomWalkwithStop> a; return run_app(file, N, R, L, source, same, p, m, a); }
    if (prog == "prob") { CountingWalks<RandomWalkwithProb> a; return run_app(file, N, R, L, source, same, p, m, a); }
    if (prog == "joint") { CountingWalks<RandomWalkwithRestartwithJoint> a; return run_app(file, N, R, L, source, same, p, m, a); }
    CountingWalks<RandomWalkwithRestart> a;
    return run_app(file, N, R, L, source, same, p, m, a);
}
