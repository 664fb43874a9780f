// walks/randomwalkwithstop.hpp -- walks that die on a dead end or with
// probability walkprob, capped at L steps
// (src/walks/randomwalkwithstop.hpp:20-52; device kind GW_STOP).
#ifndef RANDOMWALKWITHSTOP
#define RANDOMWALKWITHSTOP

#include "walks/randomwalk.hpp"

class RandomWalkwithStop : public RandomWalk {
public:
    gw_device_walk device_program() const override { return {GW_STOP, L, walkprob, 0}; }
};

#endif
