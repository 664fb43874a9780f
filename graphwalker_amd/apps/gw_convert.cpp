// gw_convert.cpp -- preprocesses a text edge list into the graph file format
// the engine loads (convert_if_notexists, preprocess/conversions.hpp):
//
//   gw_convert file <edge list> blocksize_kb <KB> [keepselfloops 1]
//
// Prints the number of vertices, edges and blocks.
#include <iostream>

#include "api/cmdopts.hpp"
#include "preprocess/conversions.hpp"
#include "util/toplist.hpp"

int main(int argc, const char** argv) {
    set_argc(argc, argv);
    const std::string file = get_option_string("file");
    const unsigned long long blocksize_kb = get_option_long("blocksize_kb", 1024);
    const bid_t nblocks = convert_if_notexists(file, blocksize_kb);
    std::cout << "nvertices " << get_num_vertices(file) << " nblocks " << nblocks << std::endl;
    return 0;
}
