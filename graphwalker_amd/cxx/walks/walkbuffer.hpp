// walks/walkbuffer.hpp -- a growable per-(thread, block) buffer of packed walk
// records with the reference's public shape (src/walks/walkbuffer.hpp:8-40):
// `size_w`, `walks`, operator[], push_back().  It grows in host memory
// instead of spilling to a pool file when WALK_BUFFER_SIZE records are held.
#ifndef GW_CXX_WALKBUFFER_HPP
#define GW_CXX_WALKBUFFER_HPP

#include <cstdlib>
#include <cstring>

#include "api/datatype.hpp"

class WalkBuffer {
public:
    bool malloced;
    wid_t size_w;
    WalkDataType* walks;

    WalkBuffer() : malloced(false), size_w(0), walks(nullptr), cap_(0) {}
    ~WalkBuffer() { std::free(walks); }
    WalkBuffer(const WalkBuffer&) = delete;
    WalkBuffer& operator=(const WalkBuffer&) = delete;

    WalkDataType& operator[](size_t i) { return walks[i]; }

    void reserve(wid_t n) {
        if (n <= cap_) return;
        WalkDataType* p = static_cast<WalkDataType*>(std::realloc(walks, n * sizeof(WalkDataType)));
        if (!p) logstream(LOG_FATAL) << "WalkBuffer: out of host memory" << std::endl;
        walks = p;
        cap_ = n;
        malloced = true;
    }
    void push_back(WalkDataType w) {
        if (size_w == cap_) reserve(cap_ ? 2 * cap_ : WALK_BUFFER_SIZE);
        walks[size_w++] = w;
    }
    wid_t size() const { return size_w; }
    void clear() { size_w = 0; }

private:
    wid_t cap_;
};

#endif
