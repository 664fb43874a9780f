// metrics/reps/file_reporter.hpp -- writes every metric to a text file and
// appends one statistics line: the run time, then the phase timers in key
// order (the reference's file reporter and its .statistics line,
// src/metrics/reps/file_reporter.hpp:66-91).
#ifndef GW_CXX_FILE_REPORTER_HPP
#define GW_CXX_FILE_REPORTER_HPP

#include <fstream>

#include "metrics/metrics.hpp"

class file_reporter : public imetrics_reporter {
public:
    explicit file_reporter(std::string fname) : fname_(fname) {}
    void do_report(std::string name, std::string id, std::map<std::string, metrics_entry>& entries) override {
        std::ofstream f(fname_.c_str(), std::ios::app);
        f << "[application]\nname=" << name << "\nid=" << id << "\n";
        for (auto& kv : entries) {
            const metrics_entry& e = kv.second;
            f << kv.first << "=";
            if (e.valtype == STRING) f << e.stringval;
            else if (e.valtype == TIME) f << e.sum;
            else f << e.value;
            f << "\n";
        }
        std::string file = entries.count("file") ? entries["file"].stringval : name;
        std::ofstream st((file + ".statistics").c_str(), std::ios::app);
        for (auto& kv : entries)
            if (kv.second.valtype == TIME) st << kv.second.sum << "\t";
        st << "\n";
    }

private:
    std::string fname_;
};

#endif
