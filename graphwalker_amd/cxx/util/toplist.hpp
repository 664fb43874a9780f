// util/toplist.hpp -- top-K vertices of a vertex-value file
// (<g>_GraphWalker/<bytes>B.vvalue), the helper apps include from
// src/util/toplist.hpp:41-183.
#ifndef DEF_GRAPHCHI_TOPLIST
#define DEF_GRAPHCHI_TOPLIST

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <sstream>
#include <string>
#include <vector>

#include "api/cmdopts.hpp"
#include "api/datatype.hpp"
#include "api/filename.hpp"
#include "api/io.hpp"

template <typename T>
struct vertex_value {
    vid_t vertex;
    T value;
    vertex_value() {}
    vertex_value(vid_t v, T x) : vertex(v), value(x) {}
};

static VARIABLE_IS_NOT_USED std::string filename_vertex_data(std::string basefilename) {
    std::ostringstream s;
    s << basefilename << "_GraphWalker/" << sizeof(VertexDataType) << "B.vvalue";
    return s.str();
}

template <typename T>
bool vertex_value_greater(const vertex_value<T>& a, const vertex_value<T>& b) {
    return a.value > b.value || (a.value == b.value && a.vertex < b.vertex);
}

static VARIABLE_IS_NOT_USED size_t get_num_vertices(std::string basefilename) {
    std::ifstream in(nverticesname(basefilename).c_str());
    size_t n = 0;
    in >> n;
    return n;
}

template <typename T>
std::vector<vertex_value<T>> get_top_vertices(std::string basefilename, int ntop, vid_t from = 0, vid_t to = 0) {
    const std::string f = filename_vertex_data(basefilename);
    int fd = open(f.c_str(), O_RDONLY);
    if (fd < 0) logstream(LOG_FATAL) << "Could not open " << f << std::endl;
    const size_t n = static_cast<size_t>(lseek(fd, 0, SEEK_END)) / sizeof(T);
    if (to == 0 || to > n) to = static_cast<vid_t>(n);
    std::vector<T> vals(to > from ? to - from : 0);
    if (!vals.empty()) preada(fd, vals.data(), vals.size() * sizeof(T), static_cast<size_t>(from) * sizeof(T));
    close(fd);
    std::vector<vertex_value<T>> all(vals.size());
    for (size_t i = 0; i < vals.size(); ++i) all[i] = vertex_value<T>(static_cast<vid_t>(from + i), vals[i]);
    const size_t k = std::min<size_t>(static_cast<size_t>(std::max(ntop, 0)), all.size());
    std::partial_sort(all.begin(), all.begin() + k, all.end(), vertex_value_greater<T>);
    all.resize(k);
    return all;
}

#endif
