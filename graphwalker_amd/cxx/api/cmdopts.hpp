// api/cmdopts.hpp -- command-line options with the reference's lookup rules
// (src/api/cmdopts.hpp:86-218): set_argc() loads the config files, then
// "--key=value" arguments; get_option_*(name, default) looks for a positional
// "name value" pair (the last occurrence wins), then the config map, then the
// default.
#ifndef GW_CXX_CMDOPTS_HPP
#define GW_CXX_CMDOPTS_HPP

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <iostream>
#include <map>
#include <string>

#include "api/configfile.hpp"
#include "api/filename.hpp"
#include "logger/logger.hpp"

#ifdef __GNUC__
#define VARIABLE_IS_NOT_USED __attribute__((unused))
#else
#define VARIABLE_IS_NOT_USED
#endif

namespace gw_opts {
struct state {
    bool configured = false;
    int argc = 0;
    char** argv = nullptr;
    std::map<std::string, std::string> conf;
};
inline state& st() {
    static state s;
    return s;
}
inline const char* positional(const char* name) {
    state& s = st();
    if (!s.configured) {
        std::cerr << "ERROR: command line options not initialized; call set_argc() first." << std::endl;
        return nullptr;
    }
    for (int i = s.argc - 2; i >= 0; --i)
        if (strcmp(s.argv[i], name) == 0) return s.argv[i + 1];
    return nullptr;
}
inline const std::string* conf(const char* name) {
    auto& c = st().conf;
    auto it = c.find(name);
    return it == c.end() ? nullptr : &it->second;
}
}  // namespace gw_opts

static void VARIABLE_IS_NOT_USED set_conf(std::string key, std::string value) { gw_opts::st().conf[key] = value; }

static void VARIABLE_IS_NOT_USED set_argc(int argc, const char** argv) {
    gw_opts::state& s = gw_opts::st();
    s.argc = argc;
    s.argv = const_cast<char**>(argv);
    s.configured = true;
    s.conf = loadconfig(configlocalname(), configname());
    for (int i = 1; i < argc; ++i) {
        std::string a(argv[i]);
        if (a.compare(0, 2, "--") != 0) continue;
        a = a.substr(2);
        const size_t eq = a.find('=');
        if (eq == std::string::npos) continue;
        s.conf[a.substr(0, eq)] = a.substr(eq + 1);
    }
}

static std::string VARIABLE_IS_NOT_USED get_option_string(const char* name, std::string dflt) {
    if (const char* p = gw_opts::positional(name)) return p;
    if (const std::string* c = gw_opts::conf(name)) return *c;
    return dflt;
}

static std::string VARIABLE_IS_NOT_USED get_option_string(const char* name) {
    if (const char* p = gw_opts::positional(name)) return p;
    if (const std::string* c = gw_opts::conf(name)) return *c;
    logstream(LOG_FATAL) << "missing required option: " << name << std::endl;
    return "";
}

static int VARIABLE_IS_NOT_USED get_option_int(const char* name, int dflt) {
    if (const char* p = gw_opts::positional(name)) return atoi(p);
    if (const std::string* c = gw_opts::conf(name)) return atoi(c->c_str());
    return dflt;
}

static int VARIABLE_IS_NOT_USED get_option_int(const char* name) {
    if (const char* p = gw_opts::positional(name)) return atoi(p);
    if (const std::string* c = gw_opts::conf(name)) return atoi(c->c_str());
    logstream(LOG_FATAL) << "missing required option: " << name << std::endl;
    return 0;
}

static uint64_t VARIABLE_IS_NOT_USED get_option_long(const char* name, uint64_t dflt) {
    if (const char* p = gw_opts::positional(name)) return strtoull(p, nullptr, 10);
    if (const std::string* c = gw_opts::conf(name)) return strtoull(c->c_str(), nullptr, 10);
    return dflt;
}

static float VARIABLE_IS_NOT_USED get_option_float(const char* name, float dflt) {
    if (const char* p = gw_opts::positional(name)) return static_cast<float>(atof(p));
    if (const std::string* c = gw_opts::conf(name)) return static_cast<float>(atof(c->c_str()));
    return dflt;
}

#endif
