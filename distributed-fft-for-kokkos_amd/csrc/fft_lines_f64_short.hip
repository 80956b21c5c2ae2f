// Instantiation unit of the compile-time-length line-pass kernels: FP64
// lengths 2, 4, 8, 16, 32, 64, 96, 128, 160 (the units are compiled in parallel).
#include "fft_line_impl.hpp"

namespace dfft {

hipError_t dispatch_fast_f64_short(const LinePass& p, hipStream_t s) {
  using Real = double;
  switch (p.n) {
    case 2: return launch_fast_r<Real, 2>(p, s);
    case 4: return launch_fast_r<Real, 4>(p, s);
    case 8: return launch_fast_r<Real, 8>(p, s);
    case 16: return launch_fast_r<Real, 16>(p, s);
    case 32: return launch_fast_r<Real, 32>(p, s);
    case 64: return launch_fast_r<Real, 64>(p, s);
    case 96: return launch_fast_r<Real, 96>(p, s);
    case 128: return launch_fast_r<Real, 128>(p, s);
    case 160: return launch_fast_r<Real, 160>(p, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace dfft
