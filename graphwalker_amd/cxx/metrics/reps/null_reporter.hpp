// metrics/reps/null_reporter.hpp -- discards metrics.
#ifndef GW_CXX_NULL_REPORTER_HPP
#define GW_CXX_NULL_REPORTER_HPP

#include "metrics/metrics.hpp"

class null_reporter : public imetrics_reporter {
public:
    void do_report(std::string, std::string, std::map<std::string, metrics_entry>&) override {}
};

#endif
