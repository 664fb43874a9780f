// api/datatype.hpp -- scalar types and sizing constants of the walk API
// (same names and widths as src/api/datatype.hpp:16-23 of GraphWalker, so
// apps declare their arrays and walk records unchanged).
#ifndef GW_CXX_DATATYPE_HPP
#define GW_CXX_DATATYPE_HPP

#include <stdint.h>
#include <vector>

#include "logger/logger.hpp"

#define FILE_SIZE 1024                   // GB: one CSR file holds the whole graph
#define VERT_SIZE (64 * 1024 * 1024)     // vertices per beg_pos buffer when converting
#define EDGE_SIZE (256 * 1024 * 1024)    // edges per csr buffer when converting
#define WALK_BUFFER_SIZE (4 * 1024)      // initial capacity of a per-(thread, block) walk buffer
#define MEM_BUDGET (44 * 1024 * 1024)    // KB of host memory for in-memory blocks (compNmblocks)

typedef uint32_t vid_t;   // vertex id
typedef uint64_t eid_t;   // edge position
typedef uint64_t wid_t;   // walk count / walk index
typedef uint32_t bid_t;   // block id
typedef uint16_t hid_t;   // hop
typedef uint8_t tid_t;    // thread id
typedef unsigned VertexDataType;
typedef unsigned long WalkDataType;  // packed walk record

#endif
