// Named ranges on the host timeline for the profiler (rocprofv3 --marker-trace): each entry point
// that launches device work brackets it with a range, so a trace shows which call a kernel
// belongs to. Without a profiler attached a range costs a function call.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace cqx {

struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

} // namespace cqx
