// roctx ranges around set-up, time stepping and stage launches, visible with
// `rocprofv3 --marker-trace` (the reference times set-up and run with
// std::chrono under SHOW_TIMER_INFORMATION, Solver.cpp:486-531,
// src/CMakeLists.txt:23-25).  A range costs a function call when no tool is
// attached.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace dgtd {

struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace dgtd
