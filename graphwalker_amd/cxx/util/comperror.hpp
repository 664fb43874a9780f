// util/comperror.hpp -- vertex-value file helpers and the top-N relative
// error against an "accurate" file (the helpers of src/util/comperror.hpp:7-97):
// initialVertexValue<T>(N, g) writes N zero values; computeError<T>(N, g,
// ntop, app) compares the top-N of g's vertex values, normalised, with
// <g>_CompError/accurate_<app>_top100.value (lines "vertex value").
#ifndef DEF_GRAPHCHI_COMPERROR
#define DEF_GRAPHCHI_COMPERROR

#include <fcntl.h>
#include <math.h>
#include <unistd.h>

#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "util/toplist.hpp"

template <typename T>
void initialVertexValue(vid_t N, std::string basefilename) {
    std::vector<T> zero(N, T(0));
    const std::string f = filename_vertex_data(basefilename);
    int fd = open(f.c_str(), O_WRONLY | O_CREAT | O_TRUNC, S_IRUSR | S_IWUSR | S_IRGRP | S_IROTH);
    if (fd < 0) logstream(LOG_FATAL) << "Could not create " << f << std::endl;
    if (N) pwritea(fd, zero.data(), zero.size() * sizeof(T), 0);
    close(fd);
}

template <typename T>
void computeError(unsigned N, std::string basefilename, int ntop, std::string app) {
    std::vector<vertex_value<T>> top = get_top_vertices<T>(basefilename, ntop);
    double total = 0;
    {
        std::vector<vertex_value<T>> all = get_top_vertices<T>(basefilename, static_cast<int>(N));
        for (auto& x : all) total += static_cast<double>(x.value);
    }
    std::ifstream acc((basefilename + "_CompError/accurate_" + app + "_top100.value").c_str());
    std::map<vid_t, double> exact;
    vid_t v;
    double x;
    while (acc >> v >> x) exact[v] = x;
    double err = 0;
    int cnt = 0;
    for (auto& t : top) {
        auto it = exact.find(t.vertex);
        if (it == exact.end() || it->second == 0 || total == 0) continue;
        err += fabs(static_cast<double>(t.value) / total - it->second) / it->second;
        ++cnt;
    }
    logstream(LOG_INFO) << "Error of top " << ntop << " (" << app << ") = " << (cnt ? err / cnt : 0.0) << std::endl;
}

#endif
