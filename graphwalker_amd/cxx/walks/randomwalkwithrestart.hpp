// walks/randomwalkwithrestart.hpp -- walks of nsteps steps that return to
// their source vertex on a dead end or with probability walkprob: the
// restart-PPR semantics of src/walks/randomwalkwithrestart.hpp:33-63 (stale in
// the reference, which no longer compiles) on the current engine API
// (device kind GW_RESTART).  The record's source field holds the source
// vertex.
#ifndef RANDOMWALKWITHRESTART
#define RANDOMWALKWITHRESTART

#include "walks/randomwalk.hpp"

class RandomWalkwithRestart : public RandomWalk {
public:
    vid_t source = 0;
    wid_t nwalks = 0;
    hid_t nsteps = 0;
    float tail = 0;
    wid_t tailwalknum = 0;

    void initializeRW(vid_t _source, wid_t _nwalks, hid_t _nsteps, float _tail) {
        source = _source;
        nwalks = _nwalks;
        nsteps = _nsteps;
        tail = _tail;
        tailwalknum = static_cast<wid_t>(nwalks * tail);
        R = nwalks;
        L = nsteps;
    }
    void initializeRW(wid_t _R, hid_t _L) override {
        R = nwalks = _R;
        L = nsteps = _L;
    }
    gw_device_walk device_program() const override { return {GW_RESTART, L, walkprob, 0}; }
};

#endif
