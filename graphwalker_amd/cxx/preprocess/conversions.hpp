// preprocess/conversions.hpp -- the preprocessed-graph file format and the
// conversion entry point apps call (src/preprocess/conversions.hpp:181-404).
//
// convert_if_notexists(g, blocksize_kb) returns the number of blocks of the
// graph `g`, creating what is missing:
//   * the CSR files <g>_GraphWalker/graphinfo/file_0.beg_pos (uint64, n+1
//     absolute offsets) and file_0.csr (uint32 columns), plus N.nvertices and
//     filesize_1024GB.filerange, from the text edge list `g` ("from to" per
//     line, '#' / '%' comments).  Rows keep the edge order of the file, edges
//     are grouped by source; self-loops are dropped like the reference does,
//     unless the option `keepselfloops 1` is given;
//   * the block table blocksize_<KB>KB.blockrange by the reference's cut rule:
//     a block grows while its edges fit blocksize_kb KB of uint32 columns; a
//     vertex whose row alone exceeds that gets a block of its own.
// gw_write_csr() writes a CSR built elsewhere in the same format.
#ifndef GW_CXX_CONVERSIONS_HPP
#define GW_CXX_CONVERSIONS_HPP

#include <errno.h>
#include <fcntl.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "api/cmdopts.hpp"
#include "api/datatype.hpp"
#include "api/filename.hpp"
#include "api/io.hpp"
#include "logger/logger.hpp"

namespace gw_conv {

inline bool exists(const std::string& f) {
    struct stat st;
    return stat(f.c_str(), &st) == 0;
}

inline void ensure_dirs(const std::string& base) {
    mkdir((base + "_GraphWalker/").c_str(), 0777);
    mkdir((base + "_GraphWalker/graphinfo/").c_str(), 0777);
}

inline size_t count_lines(const std::string& f) {
    std::ifstream in(f.c_str());
    size_t n = 0;
    std::string line;
    while (std::getline(in, line)) ++n;
    return n;
}

inline void write_binary(const std::string& f, const void* data, size_t bytes) {
    int fd = open(f.c_str(), O_WRONLY | O_CREAT | O_TRUNC, S_IRUSR | S_IWUSR | S_IRGRP | S_IROTH);
    if (fd < 0) logstream(LOG_FATAL) << "cannot create " << f << ": " << strerror(errno) << std::endl;
    pwritea(fd, static_cast<const char*>(data), bytes, 0);
    close(fd);
}

}  // namespace gw_conv

// Writes a CSR (off[n+1] absolute, col[off[n]]) in the preprocessed format.
static inline void gw_write_csr(const std::string& base, const eid_t* off, const vid_t* col, vid_t n) {
    gw_conv::ensure_dirs(base);
    const std::string f = fidname(base, 0);
    gw_conv::write_binary(f + ".beg_pos", off, (static_cast<size_t>(n) + 1) * sizeof(eid_t));
    gw_conv::write_binary(f + ".csr", col, static_cast<size_t>(off[n]) * sizeof(vid_t));
    std::ofstream fr(filerangename(base, FILE_SIZE).c_str());
    fr << 0 << "\n" << n << "\n";
    std::ofstream nv(nverticesname(base).c_str());
    nv << n << "\n";
}

static inline bid_t find_filerange(std::string base, uint16_t filesize_GB) {
    const std::string f = filerangename(base, filesize_GB);
    if (!gw_conv::exists(f)) return 0;
    const size_t lines = gw_conv::count_lines(f);
    return lines > 1 ? static_cast<bid_t>(lines - 1) : 0;
}

static inline bid_t find_blockrange(std::string base, unsigned long long blocksize_kb) {
    const std::string f = blockrangename(base, blocksize_kb);
    if (!gw_conv::exists(f)) return 0;
    const size_t lines = gw_conv::count_lines(f);
    return lines > 1 ? static_cast<bid_t>(lines - 1) : 0;
}

static inline bid_t convert_to_csr(std::string base, uint16_t /*filesize_GB*/) {
    FILE* in = fopen(base.c_str(), "r");
    if (!in) logstream(LOG_FATAL) << "Could not load :" << base << " error: " << strerror(errno) << std::endl;
    const bool keep_loops = get_option_int("keepselfloops", 0) != 0;
    std::vector<std::pair<vid_t, vid_t>> edges;
    vid_t max_vert = 0;
    char line[1024];
    while (fgets(line, sizeof line, in)) {
        if (line[0] == '#' || line[0] == '%') continue;
        char* t1 = strtok(line, "\t, ");
        char* t2 = strtok(nullptr, "\t, \r\n");
        if (!t1 || !t2) continue;
        const vid_t a = static_cast<vid_t>(strtoul(t1, nullptr, 10));
        const vid_t b = static_cast<vid_t>(strtoul(t2, nullptr, 10));
        if (a == b && !keep_loops) continue;
        max_vert = std::max(max_vert, std::max(a, b));
        edges.emplace_back(a, b);
    }
    fclose(in);
    std::stable_sort(edges.begin(), edges.end(),
                     [](const std::pair<vid_t, vid_t>& x, const std::pair<vid_t, vid_t>& y) { return x.first < y.first; });
    const vid_t n = edges.empty() ? 1 : max_vert + 1;
    std::vector<eid_t> off(static_cast<size_t>(n) + 1, 0);
    std::vector<vid_t> col(edges.size());
    for (const auto& e : edges) off[e.first + 1]++;
    for (vid_t v = 0; v < n; ++v) off[v + 1] += off[v];
    for (size_t i = 0; i < edges.size(); ++i) col[i] = edges[i].second;
    gw_write_csr(base, off.data(), col.data(), n);
    logstream(LOG_INFO) << "converted " << base << ": nverts = " << n << ", nedges = " << edges.size() << std::endl;
    return 1;
}

// The block cut rule of the reference (conversions.hpp:300-364) over the
// whole beg_pos array: returns the block table (first vertex of each block,
// then n).
static inline std::vector<vid_t> block_table(const std::vector<eid_t>& beg_pos, unsigned long long blocksize_kb) {
    const eid_t mneb = static_cast<eid_t>(blocksize_kb) * 1024 / sizeof(vid_t);
    std::vector<vid_t> blocks{0};
    eid_t start_pos = 0;
    const size_t ttv = beg_pos.size();
    for (size_t v = 1; v < ttv; ++v) {
        if (beg_pos[v] - start_pos > mneb) {
            if (beg_pos[v] - beg_pos[v - 1] > mneb) {
                blocks.push_back(static_cast<vid_t>(v - 1));
                blocks.push_back(static_cast<vid_t>(v));
                start_pos = beg_pos[v];
            } else {
                blocks.push_back(static_cast<vid_t>(v - 1));
                start_pos = beg_pos[v - 1];
            }
        }
    }
    blocks.push_back(static_cast<vid_t>(ttv - 1));
    return blocks;
}

static inline bid_t compute_block(std::string base, unsigned long long blocksize_kb) {
    const std::string bp = fidname(base, 0) + ".beg_pos";
    int fd = open(bp.c_str(), O_RDONLY);
    if (fd < 0) logstream(LOG_FATAL) << "Could not load :" << bp << ", error: " << strerror(errno) << std::endl;
    const off_t bytes = lseek(fd, 0, SEEK_END);
    std::vector<eid_t> beg_pos(static_cast<size_t>(bytes) / sizeof(eid_t));
    preada(fd, beg_pos.data(), beg_pos.size() * sizeof(eid_t), 0);
    close(fd);
    const std::vector<vid_t> blocks = block_table(beg_pos, blocksize_kb);
    std::ofstream brf(blockrangename(base, blocksize_kb).c_str());
    for (vid_t b : blocks) brf << b << "\n";
    return static_cast<bid_t>(blocks.size() - 1);
}

static inline bid_t convert_if_notexists(std::string base, unsigned long long blocksize_kb) {
    bid_t nfiles = find_filerange(base, FILE_SIZE);
    if (nfiles > 0) {
        logstream(LOG_INFO) << "Found preprocessed files for " << base << ", num file=" << nfiles << std::endl;
    } else {
        logstream(LOG_INFO) << "Did not find preprocessed files for " << base << "; converting" << std::endl;
        nfiles = convert_to_csr(base, FILE_SIZE);
    }
    if (blocksize_kb == 0) logstream(LOG_FATAL) << "blocksize_kb must be > 0" << std::endl;
    bid_t nblocks = find_blockrange(base, blocksize_kb);
    if (nblocks > 0) {
        logstream(LOG_INFO) << "Found computed blocks for " << base << ", blocksize = " << blocksize_kb
                            << "KB, num blocks=" << nblocks << std::endl;
    } else {
        nblocks = compute_block(base, blocksize_kb);
        logstream(LOG_INFO) << "computed " << nblocks << " blocks." << std::endl;
    }
    return nblocks;
}

#endif
