// api/configfile.hpp -- "key = value" configuration files (the role of
// loadconfig in src/api/configfile.hpp:81-117): the first of the two files
// that exists is read; lines starting with '#' or '%' are comments; key and
// value are the first two '='-separated tokens, trimmed.  Unlike the
// reference, a missing file yields an empty configuration instead of an
// assertion failure.
#ifndef GW_CXX_CONFIGFILE_HPP
#define GW_CXX_CONFIGFILE_HPP

#include <fstream>
#include <map>
#include <string>

namespace gw_conffile {

inline std::string trim(const std::string& s) {
    const char* ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}

inline bool read_into(const std::string& file, std::map<std::string, std::string>& out) {
    std::ifstream in(file.c_str());
    if (!in.good()) return false;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#' || line[0] == '%') continue;
        // first two tokens separated by runs of '='
        size_t a = line.find_first_not_of('=');
        if (a == std::string::npos) continue;
        size_t b = line.find('=', a);
        if (b == std::string::npos) continue;
        const std::string key = trim(line.substr(a, b - a));
        size_t c = line.find_first_not_of('=', b);
        if (c == std::string::npos) continue;
        size_t d = line.find('=', c);
        const std::string val = trim(line.substr(c, d == std::string::npos ? std::string::npos : d - c));
        if (!key.empty()) out[key] = val;
    }
    return true;
}

}  // namespace gw_conffile

static inline std::map<std::string, std::string> loadconfig(std::string filename, std::string secondary_filename) {
    std::map<std::string, std::string> conf;
    if (!gw_conffile::read_into(filename, conf)) gw_conffile::read_into(secondary_filename, conf);
    return conf;
}

#endif
