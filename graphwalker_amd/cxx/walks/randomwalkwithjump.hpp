// walks/randomwalkwithjump.hpp -- walks of exactly L steps that teleport to a
// uniform vertex on a dead end or with probability walkprob
// (src/walks/randomwalkwithjump.hpp:30-62; device kind GW_JUMP).
#ifndef RANDOMWALKWITHJUMP
#define RANDOMWALKWITHJUMP

#include "walks/randomwalk.hpp"

class RandomWalkwithJump : public RandomWalk {
public:
    vid_t N = 0;

    void initializeRW(vid_t _N, wid_t _R, hid_t _L) {
        N = _N;
        R = _R;
        L = _L;
    }
    gw_device_walk device_program() const override { return {GW_JUMP, L, walkprob, N}; }
};

#endif
