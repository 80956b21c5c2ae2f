// Host-side checks of caller buffers at the C boundary.  The reference
// compares every view's extents with the plan's before it runs
// (KokkosFFT_Distributed_InternalPlan.hpp:109-142, TransBlock.hpp:120-186);
// here a rank's array is a bare device pointer, so the check is that the
// allocation holding it extends far enough past it for the local array.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "planner.hpp"

namespace dfft {

// Throws DFFT_ERROR_MISMATCH when `p` lies in a HIP allocation that ends
// less than `bytes` past it.  Memory the runtime does not track (for
// example memory mapped by another allocator) is let through: its extent is
// unknown.  A null pointer with bytes > 0 is an invalid argument.
inline void check_device_extent(const void* p, std::size_t bytes, const std::string& what) {
  if (bytes == 0) return;
  if (p == nullptr) throw Error(DFFT_ERROR_INVALID_ARGUMENT, what + ": null pointer for a non-empty array");
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  const char* b = static_cast<const char*>(base);
  const char* q = static_cast<const char*>(p);
  if (q + bytes > b + size)
    throw Error(DFFT_ERROR_MISMATCH, what + ": the allocation holds " + std::to_string(b + size - q) +
                                         " bytes past the pointer, the local array needs " + std::to_string(bytes));
}

}  // namespace dfft
