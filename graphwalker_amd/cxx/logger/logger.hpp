// logger/logger.hpp -- leveled line logger with the call shape the walk apps
// use: logstream(LOG_INFO) << ... << std::endl (src/logger/logger.hpp:71-116,
// 215-216).  A line at LOG_FATAL throws when it ends, as the reference does.
// The threshold comes from the environment variable GW_LOG_LEVEL (0 = debug
// ... 5 = none; default 1 = info).
#ifndef GW_CXX_LOGGER_HPP
#define GW_CXX_LOGGER_HPP

#include <cstdlib>
#include <iostream>
#include <mutex>
#include <sstream>
#include <string>

#define LOG_DEBUG 0
#define LOG_INFO 1
#define LOG_WARNING 2
#define LOG_ERROR 3
#define LOG_FATAL 4
#define LOG_NONE 5

namespace gw_log {

inline int threshold() {
    static int t = [] {
        const char* e = std::getenv("GW_LOG_LEVEL");
        return e ? std::atoi(e) : LOG_INFO;
    }();
    return t;
}

inline std::mutex& out_mutex() {
    static std::mutex m;
    return m;
}

class line {
public:
    explicit line(int level) : level_(level) {}
    line(const line&) = delete;
    line(line&& o) noexcept : level_(o.level_), buf_(std::move(o.buf_)) { o.level_ = -1; }
    ~line() { flush(false); }

    template <class T>
    line& operator<<(const T& x) {
        buf_ << x;
        return *this;
    }
    // std::endl and other manipulators end the line
    line& operator<<(std::ostream& (*manip)(std::ostream&)) {
        if (manip == static_cast<std::ostream& (*)(std::ostream&)>(std::endl)) {
            flush(true);
        } else {
            buf_ << manip;
        }
        return *this;
    }

private:
    void flush(bool ended) {
        if (level_ < 0) return;
        const std::string s = buf_.str();
        if (!s.empty() && level_ >= threshold() && level_ < LOG_NONE) {
            static const char* names[] = {"DEBUG", "INFO", "WARNING", "ERROR", "FATAL"};
            std::lock_guard<std::mutex> g(out_mutex());
            std::cerr << names[level_] << ":  " << s << std::endl;
        }
        buf_.str("");
        buf_.clear();
        const int lvl = level_;
        if (ended && lvl == LOG_FATAL) {
            level_ = -1;
            throw "log fatal";
        }
    }
    int level_;
    std::ostringstream buf_;
};

}  // namespace gw_log

inline gw_log::line logstream(int level) { return gw_log::line(level); }

#endif
