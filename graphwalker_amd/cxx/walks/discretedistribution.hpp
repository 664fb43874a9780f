// walks/discretedistribution.hpp -- the bounded per-source visit histogram
// msppr records into (src/walks/discretedistribution.hpp:7-95): at most
// `capacity` (1000) distinct vertices with 16-bit counts; when full, entries
// below a minimum count are dropped and the survivors' counts reduced by
// mincount - 1.  getTop(k) logs the k most visited vertices.
#ifndef DEF_DISCRETE_DISTRIBUTION
#define DEF_DISCRETE_DISTRIBUTION

#include <stdint.h>

#include <algorithm>
#include <queue>
#include <vector>

#include "api/datatype.hpp"
#include "logger/logger.hpp"

const unsigned capacity = 1000;

struct IdCount {
    vid_t id;
    uint16_t count;
    IdCount(vid_t i, uint16_t c) : id(i), count(c) {}
    bool operator<(const IdCount o) const { return count > o.count; }
};

class DiscreteDistribution {
public:
    std::vector<vid_t> ids;
    std::vector<uint16_t> counts;
    unsigned size;

    DiscreteDistribution() : ids(capacity), counts(capacity), size(0) {}

    void add(vid_t id) {
        for (unsigned i = 0; i < size; ++i)
            if (ids[i] == id) {
                ++counts[i];
                return;
            }
        while (size >= capacity) filter(2);
        ids[size] = id;
        counts[size] = 1;
        ++size;
    }

    void filter(uint16_t mincount) {
        unsigned k = 0;
        for (unsigned i = 0; i < size; ++i)
            if (counts[i] >= mincount) {
                ids[k] = ids[i];
                counts[k] = static_cast<uint16_t>(counts[i] - mincount + 1);
                ++k;
            }
        size = k;
    }

    std::vector<IdCount> top(unsigned ntop) const {
        std::priority_queue<IdCount> q;  // min-heap on count
        for (unsigned i = 0; i < size; ++i) {
            q.push(IdCount(ids[i], counts[i]));
            if (q.size() > ntop) q.pop();
        }
        std::vector<IdCount> out;
        while (!q.empty()) {
            out.push_back(q.top());
            q.pop();
        }
        std::reverse(out.begin(), out.end());
        return out;
    }

    void getTop(unsigned ntop) {
        logstream(LOG_INFO) << "getTop " << ntop << " from size = " << size << std::endl;
        const std::vector<IdCount> t = top(ntop);
        logstream(LOG_INFO) << "Top " << ntop << " visitfrequencies - " << std::endl;
        // listed as the reference lists them, fewest visits first
        for (size_t i = 0; i < t.size(); ++i)
            logstream(LOG_INFO) << i << "-\t" << t[t.size() - 1 - i].id << ":\t " << t[t.size() - 1 - i].count << std::endl;
    }
};

#endif
