// roctx ranges around the host-side phases of the drivers (enqueue of the
// predictor, corrector, step, run, adjoint sweeps), visible in rocprofv3
// --marker-trace timelines next to the kernels they launch.  A range costs a
// call into the roctx library, a no-op when no tool is attached.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace exa {

struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace exa
