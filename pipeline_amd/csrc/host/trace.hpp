// trace.hpp -- named ranges around the host steps of the path, for rocprofv3
// --marker-trace.  The names are those of the reference's own profiler entries
// for the same steps (dp::util::ProfileEntry): "cull"
// (dp/culling/cpu/src/ManagerImpl.cpp:469), "ResultBitSet::updateChanged"
// (dp/culling/src/ResultBitSet.cpp:63), "cull::getBoundingBox"
// (dp/culling/src/ManagerBitSet.cpp:153) and "Update TransformTree"
// (dp/sg/xbar/src/SceneTree.cpp:157).  Without a profiler attached a range
// costs two calls into the roctx library.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace dpcull {

class TraceRange {
 public:
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace dpcull
