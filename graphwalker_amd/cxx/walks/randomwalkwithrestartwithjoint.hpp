// walks/randomwalkwithrestartwithjoint.hpp -- walks that return to their
// source on a dead end or with probability walkprob and end when
// hop % L == L - 1 (segments of L hops stored in one record's hop field,
// src/walks/randomwalkwithrestartwithjoint.hpp:29-61; device kind
// GW_RESTART_JOINT).  Like the reference, it declares its own R and L.
#ifndef RANDOMWALKWITHRESTARTWITHJOINT
#define RANDOMWALKWITHRESTARTWITHJOINT

#include "walks/randomwalk.hpp"

class RandomWalkwithRestartwithJoint : public RandomWalk {
public:
    wid_t R = 0;
    hid_t L = 0;

    void initializeRW(wid_t _R, hid_t _L) override {
        R = _R;
        L = _L;
    }
    gw_device_walk device_program() const override { return {GW_RESTART_JOINT, L, walkprob, 0}; }
};

#endif
