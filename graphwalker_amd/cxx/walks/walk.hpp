// walks/walk.hpp -- WalkManager: packed walk records and their per-block
// buckets, with the public members and methods apps use in startWalksbyApp
// (src/walks/walk.hpp:17-267): walknum, dwalknum, minstep, pwalks[t][p],
// walksum, ismodified; encode / getSourceId / getCurrentId / getHop /
// reencode / moveWalk; the block-choice heuristics.
//
// Records keep the reference's packing (source:24 | offset-in-block:26 |
// hop:14 in 64 bits).  Buckets live in host memory; graphwalker_engine::run
// gathers them into device walk records (source, global vertex, hop, id).
#ifndef GW_CXX_WALK_HPP
#define GW_CXX_WALK_HPP

#include <assert.h>
#include <stdlib.h>
#include <string.h>

#include <cstdio>
#include <string>

#include "api/datatype.hpp"
#include "api/filename.hpp"
#include "metrics/metrics.hpp"
#include "walks/walkbuffer.hpp"

class WalkManager {
protected:
    std::string base_filename;
    bid_t nblocks;
    tid_t nthreads;
    metrics& m;

public:
    wid_t* walknum;   // walks in each block
    wid_t* dwalknum;  // walks of each block spilled to disk (always 0 here)
    hid_t* minstep;
    WalkBuffer** pwalks;  // pwalks[thread][block]

    bid_t curp;
    WalkDataType* curwalks;
    wid_t walksum;
    bool* ismodified;

    WalkManager(metrics& _m, bid_t _nblocks, tid_t _nthreads, std::string _base_filename)
        : base_filename(_base_filename), nblocks(_nblocks), nthreads(_nthreads), m(_m), curp(0), curwalks(nullptr),
          walksum(0) {
        pwalks = new WalkBuffer*[nthreads];
        for (tid_t t = 0; t < nthreads; ++t) pwalks[t] = new WalkBuffer[nblocks];
        walknum = static_cast<wid_t*>(calloc(nblocks, sizeof(wid_t)));
        dwalknum = static_cast<wid_t*>(calloc(nblocks, sizeof(wid_t)));
        minstep = static_cast<hid_t*>(malloc(nblocks * sizeof(hid_t)));
        memset(minstep, 0xff, nblocks * sizeof(hid_t));
        ismodified = static_cast<bool*>(calloc(nblocks, sizeof(bool)));
    }
    ~WalkManager() {
        for (tid_t t = 0; t < nthreads; ++t) delete[] pwalks[t];
        delete[] pwalks;
        free(walknum);
        free(dwalknum);
        free(minstep);
        free(ismodified);
        free(curwalks);
    }
    WalkManager(const WalkManager&) = delete;
    WalkManager& operator=(const WalkManager&) = delete;

    bid_t num_blocks() const { return nblocks; }
    tid_t num_threads() const { return nthreads; }

    WalkDataType encode(vid_t sourceId, vid_t currentId, hid_t hop) {
        assert(hop < 16384);
        return ((static_cast<WalkDataType>(sourceId) & 0xffffff) << 40) |
               ((static_cast<WalkDataType>(currentId) & 0x3ffffff) << 14) | (static_cast<WalkDataType>(hop) & 0x3fff);
    }
    vid_t getSourceId(WalkDataType walk) { return static_cast<vid_t>(walk >> 40) & 0xffffff; }
    vid_t getCurrentId(WalkDataType walk) { return static_cast<vid_t>(walk >> 14) & 0x3ffffff; }
    hid_t getHop(WalkDataType walk) { return static_cast<hid_t>(walk & 0x3fff); }
    WalkDataType reencode(WalkDataType walk, vid_t toVertex) { return encode(getSourceId(walk), toVertex, getHop(walk)); }

    // Appends the record, re-encoded with `toVertex` as its offset in block p,
    // to thread t's bucket of block p (walk.hpp:93-101).
    void moveWalk(WalkDataType walk, bid_t p, tid_t t, vid_t toVertex) {
        assert(p < nblocks && t < nthreads);
        pwalks[t][p].push_back(reencode(walk, toVertex));
    }
    void writeWalks2Disk(tid_t, bid_t) {}  // buckets grow in memory

    // Gathers block p's bucket into curwalks (walk.hpp:114-137).
    wid_t getCurrentWalks(bid_t p) {
        wid_t count = 0;
        for (tid_t t = 0; t < nthreads; ++t) count += pwalks[t][p].size_w;
        free(curwalks);
        curwalks = static_cast<WalkDataType*>(malloc((count ? count : 1) * sizeof(WalkDataType)));
        wid_t k = 0;
        for (tid_t t = 0; t < nthreads; ++t) {
            WalkBuffer& b = pwalks[t][p];
            for (wid_t w = 0; w < b.size_w; ++w) curwalks[k++] = b[w];
            b.size_w = 0;
        }
        curp = p;
        return count;
    }
    void readWalksfromDisk(bid_t) {}

    void updateWalkNum(bid_t p) {
        wid_t forward = 0;
        for (bid_t b = 0; b < nblocks; ++b) {
            if (!ismodified[b] || b == p) continue;
            ismodified[b] = false;
            wid_t nw = dwalknum[b];
            for (tid_t t = 0; t < nthreads; ++t) nw += pwalks[t][b].size_w;
            forward += nw - walknum[b];
            walknum[b] = nw;
        }
        walksum += forward;
        walksum -= walknum[p];
        walknum[p] = 0;
        minstep[p] = 0xffff;
        free(curwalks);
        curwalks = nullptr;
    }

    void setMinStep(bid_t p, hid_t hop) {
        if (minstep[p] > hop) minstep[p] = hop;
    }
    bid_t blockWithMaxWalks() {
        wid_t maxw = 0;
        bid_t maxp = 0;
        for (bid_t p = 0; p < nblocks; ++p)
            if (maxw < walknum[p]) { maxw = walknum[p]; maxp = p; }
        return maxp;
    }
    bid_t blockWithMinStep() {
        hid_t mins = 0xffff;
        bid_t minp = 0;
        for (bid_t p = 0; p < nblocks; ++p)
            if (mins > minstep[p]) { mins = minstep[p]; minp = p; }
        return walknum[minp] > 0 ? minp : blockWithMaxWalks();
    }
    bid_t blockWithMaxWeight() {
        float maxwt = 0;
        bid_t maxp = 0;
        for (bid_t p = 0; p < nblocks; ++p) {
            const float w = minstep[p] ? static_cast<float>(walknum[p]) / minstep[p] : static_cast<float>(walknum[p]);
            if (maxwt < w) { maxwt = w; maxp = p; }
        }
        return maxp;
    }
    bid_t blockWithRandom() { return static_cast<bid_t>(rand() % nblocks); }
    bid_t chooseBlock(float prob) {
        const float cc = static_cast<float>(rand()) / static_cast<float>(RAND_MAX);
        return cc < prob ? blockWithMinStep() : blockWithMaxWalks();
    }
    void printWalksDistribution(bid_t exec_block) {
        wid_t sum = 0;
        for (bid_t p = 0; p < nblocks; ++p) sum += walknum[p];
        FILE* f = fopen((base_filename + ".walks").c_str(), "a");
        if (f) {
            fprintf(f, "%u \t %lu \t %lu\n", exec_block, static_cast<unsigned long>(walknum[exec_block]),
                    static_cast<unsigned long>(sum));
            fclose(f);
        }
    }
};

#endif
