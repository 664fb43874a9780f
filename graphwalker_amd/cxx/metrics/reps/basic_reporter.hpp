// metrics/reps/basic_reporter.hpp -- prints every metric to stdout
// (the console reporter of src/metrics/reps/basic_reporter.hpp).
#ifndef GW_CXX_BASIC_REPORTER_HPP
#define GW_CXX_BASIC_REPORTER_HPP

#include <iostream>

#include "metrics/metrics.hpp"

class basic_reporter : public imetrics_reporter {
public:
    void do_report(std::string name, std::string id, std::map<std::string, metrics_entry>& entries) override {
        std::cout << "\n=== REPORT FOR " << name << "(" << id << ") ===" << std::endl;
        for (auto& kv : entries) {
            const metrics_entry& e = kv.second;
            std::cout << "[" << kv.first << "] ";
            if (e.valtype == STRING) std::cout << e.stringval;
            else if (e.valtype == TIME) std::cout << "time=" << e.sum << "s count=" << e.count;
            else if (e.valtype == INTEGER) std::cout << static_cast<unsigned long long>(e.value);
            else std::cout << e.value;
            std::cout << std::endl;
        }
    }
};

#endif
