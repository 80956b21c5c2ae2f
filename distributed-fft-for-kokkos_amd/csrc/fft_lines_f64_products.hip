// Instantiation unit of the FP64 products passes (paired R2C with the
// divergence-form products formed at the loads): lengths 8 ... 1024.
#include "fft_line_impl.hpp"

namespace dfft {

hipError_t dispatch_products_f64(const LinePass& p, hipStream_t s) {
  switch (p.n) {
    case 8: return launch_products<8>(p, s);
    case 16: return launch_products<16>(p, s);
    case 32: return launch_products<32>(p, s);
    case 64: return launch_products<64>(p, s);
    case 128: return launch_products<128>(p, s);
    case 256: return launch_products<256>(p, s);
    case 512: return launch_products<512>(p, s);
    case 1024: return launch_products<1024>(p, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace dfft
