// api/filename.hpp -- on-disk names of a preprocessed graph.  The paths are
// the reference's file format (src/api/filename.hpp:16-76), so a graph
// preprocessed by GraphWalker loads unchanged:
//   <g>_GraphWalker/graphinfo/file_0.{beg_pos,csr}, <g>_GraphWalker/N.nvertices,
//   <g>_GraphWalker/filesize_<GB>GB.filerange, <g>_GraphWalker/blocksize_<KB>KB.blockrange
#ifndef GW_CXX_FILENAME_HPP
#define GW_CXX_FILENAME_HPP

#include <stdlib.h>

#include <sstream>
#include <string>

#include "api/datatype.hpp"

static inline std::string fidname(std::string base, bid_t p) {
    std::ostringstream s;
    s << base << "_GraphWalker/graphinfo/file_" << p;
    return s.str();
}

static inline std::string walksname(std::string base, bid_t p) {
    std::ostringstream s;
    s << base << "_GraphWalker/walks/pool_" << p << ".walks";
    return s.str();
}

static inline std::string filerangename(std::string base, uint16_t filesize_GB) {
    std::ostringstream s;
    s << base << "_GraphWalker/filesize_" << filesize_GB << "GB.filerange";
    return s.str();
}

static inline std::string blockrangename(std::string base, unsigned long long blocksize_KB) {
    std::ostringstream s;
    s << base << "_GraphWalker/blocksize_" << blocksize_KB << "KB.blockrange";
    return s.str();
}

static inline std::string nverticesname(std::string base) { return base + "_GraphWalker/N.nvertices"; }

static inline std::string configname() {
    const char* root = getenv("GRAPHCHI_ROOT");
    return root ? std::string(root) + "/conf/graphchi.cnf" : std::string("conf/graphwalker.cnf");
}

static inline std::string configlocalname() {
    const char* root = getenv("GRAPHCHI_ROOT");
    return root ? std::string(root) + "/conf/graphwalker.local.cnf" : std::string("conf/graphwalker.local.cnf");
}

#endif
