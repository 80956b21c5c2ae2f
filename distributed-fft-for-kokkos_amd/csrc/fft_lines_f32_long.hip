// Instantiation unit of the compile-time-length line-pass kernels: FP32
// lengths 256, 512, 1024, 2048 (the units are compiled in parallel).
#include "fft_line_impl.hpp"

namespace dfft {

hipError_t dispatch_fast_f32_long(const LinePass& p, hipStream_t s) {
  using Real = float;
  switch (p.n) {
    case 256: return launch_fast_r<Real, 256>(p, s);
    case 512: return launch_fast_r<Real, 512>(p, s);
    case 1024: return launch_fast_r<Real, 1024>(p, s);
    case 2048: return launch_fast_r<Real, 2048>(p, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace dfft
