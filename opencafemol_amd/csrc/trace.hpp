// trace.hpp -- named host-side ranges on the ROCTx marker stream.
//
// `rocprofv3 --marker-trace` records them beside the kernel dispatches, so a
// trace shows which step call, chunk, list rebuild, graph capture or server
// repack a kernel belongs to.  With no tool attached a range costs two calls
// into the marker library.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace ocm {

class TraceRange {
public:
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace ocm
