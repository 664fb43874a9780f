// metrics/metrics.hpp -- named wall-clock timers, counters and strings, with
// the interface the engine and apps use (src/metrics/metrics.hpp:162-321):
// start_time(key) / stop_time(key) accumulate seconds under `key`; set(key, v)
// records a value; report(reporter) hands the entries to a reporter.  Unlike
// the reference, every update is taken under a lock (its stop_time is not).
#ifndef GW_CXX_METRICS_HPP
#define GW_CXX_METRICS_HPP

#include <sys/time.h>

#include <algorithm>

#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

enum metrictype { REAL, INTEGER, TIME, STRING, VECTOR };

struct metrics_entry {
    double value = 0, sum = 0, minvalue = 0, maxvalue = 0;
    size_t count = 0;
    std::string stringval;
    metrictype valtype = REAL;
    double timer_start = -1;  // seconds, while a timer runs
    std::vector<double> v;

    metrics_entry() {}
    metrics_entry(double x, metrictype t) : value(x), sum(x), minvalue(x), maxvalue(x), count(1), valtype(t) {}
    explicit metrics_entry(const std::string& s) : count(1), stringval(s), valtype(STRING) {}
    explicit metrics_entry(metrictype t) : valtype(t) {}

    void add(double x) {
        value = x;
        sum += x;
        minvalue = count ? std::min(minvalue, x) : x;
        maxvalue = count ? std::max(maxvalue, x) : x;
        ++count;
    }
};

class imetrics_reporter {
public:
    virtual ~imetrics_reporter() {}
    virtual void do_report(std::string name, std::string id, std::map<std::string, metrics_entry>& entries) = 0;
};

class metrics {
public:
    explicit metrics(std::string name = "default", std::string id = "") : name_(name), id_(id) {}

    static double now() {
        timeval t;
        gettimeofday(&t, nullptr);
        return t.tv_sec + t.tv_usec * 1e-6;
    }

    void add(std::string key, double value, metrictype type = REAL) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = entries_.find(key);
        if (it == entries_.end()) entries_[key] = metrics_entry(value, type);
        else it->second.add(value);
    }
    void set(std::string key, size_t value) { set(key, static_cast<double>(value), INTEGER); }
    void set(std::string key, int value) { set(key, static_cast<double>(value), INTEGER); }
    void set(std::string key, double value, metrictype type = REAL) {
        std::lock_guard<std::mutex> g(mu_);
        entries_[key] = metrics_entry(value, type);
    }
    void set_integer(std::string key, size_t value) { set(key, static_cast<double>(value), INTEGER); }
    void set(std::string key, std::string s) {
        std::lock_guard<std::mutex> g(mu_);
        entries_[key] = metrics_entry(s);
    }
    void start_time(std::string key) {
        std::lock_guard<std::mutex> g(mu_);
        metrics_entry& e = entries_[key];
        e.valtype = TIME;
        e.timer_start = now();
    }
    void stop_time(std::string key) {
        const double t = now();
        std::lock_guard<std::mutex> g(mu_);
        auto it = entries_.find(key);
        if (it == entries_.end() || it->second.timer_start < 0) return;
        metrics_entry& e = it->second;
        e.add(t - e.timer_start);
        e.timer_start = -1;
    }
    double get(std::string key) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = entries_.find(key);
        return it == entries_.end() ? 0.0 : it->second.sum;
    }
    void report(imetrics_reporter& r) {
        std::lock_guard<std::mutex> g(mu_);
        r.do_report(name_, id_, entries_);
    }

private:
    std::string name_, id_;
    std::map<std::string, metrics_entry> entries_;
    std::mutex mu_;
};

#endif
