// walks/randomwalk.hpp -- base class of walk programs with the virtual
// interface apps override (src/walks/randomwalk.hpp:16-116): startWalksbyApp
// places walks, updateInfo records one visit, compUtilization / before- and
// after-block hooks.  The step semantics themselves run on the GPU: each
// program class names its kind, hop cap and stop probability through
// device_program(), and graphwalker_engine::run hands them to the device step
// kernel (libgwhip) instead of calling a per-walk updateByWalk on the host.
#ifndef RANDOMWALK
#define RANDOMWALK

#include <math.h>
#include <time.h>

#include <fstream>
#include <string>

#include "gwhip.h"
#include "api/datatype.hpp"
#include "logger/logger.hpp"
#include "walks/walk.hpp"

// What the device needs to run a program.
struct gw_device_walk {
    uint32_t kind;   // GW_JUMP .. GW_RESTART
    uint32_t L;      // hop cap, or the RESTART_JOINT period
    float p;         // stop / restart / teleport probability
    vid_t N;         // teleport range (GW_JUMP)
};

class RandomWalk {
public:
    bid_t nblocks = 0;
    vid_t* blocks = nullptr;
    wid_t R = 0;
    hid_t L = 0;
    // The reference hard-codes 0.15 in every updateByWalk (e.g.
    // randomwalkwithstop.hpp:35); kept as the default, settable per program.
    float walkprob = 0.15f;

    virtual ~RandomWalk() {}

    virtual void startWalksbyApp(WalkManager& walk_manager) {
        (void)walk_manager;
        logstream(LOG_ERROR) << "No definition of function : startWalksbyApp!" << std::endl;
    }

    virtual void updateInfo(vid_t s, vid_t dstId, tid_t threadid, hid_t hop) {
        (void)s; (void)dstId; (void)threadid; (void)hop;
        logstream(LOG_ERROR) << "No definition of function : updateInfo!" << std::endl;
    }

    // The host per-walk update of the reference (randomwalk.hpp:39).  Walks
    // step on the GPU here; this entry point is kept for source compatibility
    // and is not called by the engine.
    virtual void updateByWalk(WalkDataType walk, wid_t walkid, bid_t exec_block, eid_t*& beg_pos, vid_t*& csr,
                              WalkManager& walk_manager) {
        (void)walk; (void)walkid; (void)exec_block; (void)beg_pos; (void)csr; (void)walk_manager;
        logstream(LOG_ERROR) << "updateByWalk is not called by the device engine" << std::endl;
    }

    virtual wid_t before_exec_block(bid_t, vid_t, vid_t, WalkManager&) { return 0; }
    virtual void after_exec_block(bid_t, vid_t, vid_t, WalkManager&) {}
    virtual void compUtilization(eid_t) {}

    virtual void initializeRW(wid_t _R, hid_t _L) {
        R = _R;
        L = _L;
    }

    virtual void startWalks(WalkManager& walk_manager, bid_t _nblocks, vid_t* _blocks, std::string base_filename) {
        (void)base_filename;
        nblocks = _nblocks;
        blocks = _blocks;
        startWalksbyApp(walk_manager);
    }

    // Block of v, or nblocks when v is outside the graph (randomwalk.hpp:74-82),
    // by binary search over the block table.
    virtual unsigned getblock(vid_t v) {
        if (nblocks == 0 || v >= blocks[nblocks]) return nblocks;
        unsigned lo = 0, hi = nblocks;
        while (hi - lo > 1) {
            const unsigned mid = (lo + hi) / 2;
            if (blocks[mid] <= v) lo = mid; else hi = mid;
        }
        return lo;
    }

    virtual bool hasFinishedWalk(WalkManager& walk_manager) { return walk_manager.walksum > 0; }

    // Block size from the walk count: 2^(11 + decimal digits of nwalks) KB
    // (randomwalk.hpp:96-105).
    virtual unsigned long long compBlockSize(wid_t nwalks) {
        int digits = 0;
        while (nwalks) { ++digits; nwalks /= 10; }
        return 1ull << (11 + digits);
    }

    virtual bid_t compNmblocks(unsigned long long blocksize_kb) { return static_cast<bid_t>(MEM_BUDGET / blocksize_kb); }

    virtual gw_device_walk device_program() const = 0;
};

#endif
