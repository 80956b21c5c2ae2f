// Named host ranges around the library's steps (rocprofv3 --marker-trace
// lists them beside the kernels they launch), as the reference labels its
// exchange and pack/unpack kernels (KokkosFFT_Distributed_MPI_All2All.hpp:
// 29-32, KokkosFFT_Distributed_PackUnpack.hpp:143, 315).  Without a profiler
// attached a range costs two library calls.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdio>

namespace dfft {

class TraceRange {
 public:
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  // printf-style name (at most 127 characters)
  template <typename... Args>
  TraceRange(const char* fmt, Args... args) {
    char buf[128];
    std::snprintf(buf, sizeof(buf), fmt, args...);
    roctxRangePushA(buf);
  }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace dfft
