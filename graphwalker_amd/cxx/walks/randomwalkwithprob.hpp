// walks/randomwalkwithprob.hpp -- walks that die on a dead end or with
// probability walkprob, with no hop cap
// (src/walks/randomwalkwithprob.hpp:20-57; device kind GW_PROB).
#ifndef RANDOMWALKWITHPROB
#define RANDOMWALKWITHPROB

#include "walks/randomwalk.hpp"

class RandomWalkwithProb : public RandomWalk {
public:
    gw_device_walk device_program() const override { return {GW_PROB, L, walkprob, 0}; }
};

#endif
