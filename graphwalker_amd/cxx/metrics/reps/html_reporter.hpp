// metrics/reps/html_reporter.hpp -- metrics as an HTML table (the role of
// src/metrics/reps/html_reporter.hpp).
#ifndef GW_CXX_HTML_REPORTER_HPP
#define GW_CXX_HTML_REPORTER_HPP

#include <fstream>

#include "metrics/metrics.hpp"

class html_reporter : public imetrics_reporter {
public:
    explicit html_reporter(std::string fname) : fname_(fname) {}
    void do_report(std::string name, std::string id, std::map<std::string, metrics_entry>& entries) override {
        std::ofstream f(fname_.c_str());
        f << "<html><body><h1>" << name << " " << id << "</h1><table>\n";
        for (auto& kv : entries) {
            const metrics_entry& e = kv.second;
            f << "<tr><td>" << kv.first << "</td><td>";
            if (e.valtype == STRING) f << e.stringval;
            else if (e.valtype == TIME) f << e.sum << " s";
            else f << e.value;
            f << "</td></tr>\n";
        }
        f << "</table></body></html>\n";
    }

private:
    std::string fname_;
};

#endif
