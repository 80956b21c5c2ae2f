// Instantiation unit of the compile-time-length line-pass kernels: FP32
// lengths 4096 and 8192 (the units are compiled in parallel).
#include "fft_line_impl.hpp"

namespace dfft {

hipError_t dispatch_fast_f32_xl(const LinePass& p, hipStream_t s) {
  using Real = float;
  switch (p.n) {
    case 4096: return launch_fast_r<Real, 4096>(p, s);
    case 8192: return launch_fast_r<Real, 8192>(p, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace dfft
