// api/graphwalker_basic_includes.hpp -- the one header walk apps include
// (src/api/graphwalker_basic_includes.hpp): options, logger, metrics and
// reporters, preprocessing, the engine, plus metrics_report().
#ifndef GRAPHWALKER_DEF_ALLBASIC_INCLUDES
#define GRAPHWALKER_DEF_ALLBASIC_INCLUDES

#include <omp.h>
#include <string.h>

#include <sstream>
#include <string>

#include "api/cmdopts.hpp"
#include "api/filename.hpp"
#include "engine/graphwalker.hpp"
#include "logger/logger.hpp"
#include "metrics/metrics.hpp"
#include "metrics/reps/basic_reporter.hpp"
#include "metrics/reps/file_reporter.hpp"
#include "metrics/reps/html_reporter.hpp"
#include "preprocess/conversions.hpp"

// Reports to the console and to graphwalker_metrics.txt, like the
// reference's fixed "console,file" reporter list.
static VARIABLE_IS_NOT_USED void metrics_report(metrics& m) {
    basic_reporter console;
    m.report(console);
    file_reporter file(get_option_string("metrics.reporter.filename", "graphwalker_metrics.txt"));
    m.report(file);
}

#endif
