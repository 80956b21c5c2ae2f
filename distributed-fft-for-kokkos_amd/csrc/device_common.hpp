// Shared device-side definitions: complex helpers, fast integer division and
// the descriptor of a distributed array held as P virtual-rank buffers in one
// GPU's HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfft {

constexpr int kMaxDim = 4;     // logical rank of the arrays handled on device
constexpr int kMaxRanks = 64;  // virtual ranks per descriptor (pointer table in kernel args)

template <typename Real> struct Cplx;
template <> struct Cplx<float> { using type = float2; };
template <> struct Cplx<double> { using type = double2; };
template <typename Real> using cplx_t = typename Cplx<Real>::type;

template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }
// multiply by -i
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { return C{a.y, -a.x}; }

// Division by a runtime-constant 32-bit divisor d >= 1 for dividends below
// 2^31 (round-up multiplier method, Granlund & Montgomery, PLDI 1994):
// q = (mulhi(n, m) + n) >> l with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1.
struct FastDiv {
  uint32_t d = 1, m = 1, l = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t div) {
    d = div == 0 ? 1 : div;
    l = 0;
    while ((1ull << l) < d) ++l;
    m = static_cast<uint32_t>(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t t = __umulhi(n, m);
#else
    uint32_t t = static_cast<uint32_t>((static_cast<uint64_t>(n) * m) >> 32);
#endif
    return (t + n) >> l;
  }
};

// A logical array of `ndim` axes distributed over the ranks of a topology
// (one bin per topology factor along each axis, first ext%ntopo bins one
// element longer -- KokkosFFT_Distributed_MPI_Extents.hpp:195-235 and
// PackUnpack.hpp:40-46).  Rank r = sum_a coord[a]*rstride[a] stores its block
// as a dense local array whose storage axis k holds logical axis map[k]
// (LayoutRight storage, or LayoutLeft when layout_left is set).
struct DDesc {
  int32_t ndim;
  int32_t layout_left;
  int32_t ext[kMaxDim];
  int32_t ntopo[kMaxDim];
  int32_t rstride[kMaxDim];
  int32_t base[kMaxDim];  // ext / ntopo
  int32_t rem[kMaxDim];   // ext % ntopo
  int32_t map[kMaxDim];
  FastDiv div_big[kMaxDim];    // base + 1
  FastDiv div_small[kMaxDim];  // base
  // > 0: row pitch (elements) of the innermost storage axis of a LayoutRight
  // single-rank array, padded so rows start on 128-byte boundaries
  int32_t pitch;
  char* ptr[kMaxRanks];
};

// Owner bin and local index of global index i along axis a.
__device__ __forceinline__ void owner_local(const DDesc& d, int a, int32_t i, int32_t& c, int32_t& l,
                                            int32_t& len) {
  if (d.ntopo[a] == 1) {
    c = 0;
    l = i;
    len = d.ext[a];
    return;
  }
  const int32_t b = d.base[a], r = d.rem[a];
  const int32_t split = r * (b + 1);
  if (i < split) {
    c = static_cast<int32_t>(d.div_big[a].div(static_cast<uint32_t>(i)));
    l = i - c * (b + 1);
    len = b + 1;
  } else {
    int32_t q = static_cast<int32_t>(d.div_small[a].div(static_cast<uint32_t>(i - split)));
    c = r + q;
    l = i - split - q * b;
    len = b;
  }
}

// Byte offset of the global multi-index `idx` in rank-buffer `rank` of d
// (element size esize); the stride of axis `sa` (in elements) is returned too.
__device__ __forceinline__ char* element_ptr(const DDesc& d, const int32_t* idx, int64_t esize, int sa,
                                             int64_t& stride_sa) {
  int32_t lidx[kMaxDim], llen[kMaxDim];
  int32_t rank = 0;
#pragma unroll
  for (int a = 0; a < kMaxDim; ++a) {
    if (a < d.ndim) {
      int32_t c;
      owner_local(d, a, idx[a], c, lidx[a], llen[a]);
      rank += c * d.rstride[a];
    }
  }
  int64_t off = 0, acc = 1;
  stride_sa = 0;
  if (!d.layout_left) {
#pragma unroll
    for (int k = kMaxDim - 1; k >= 0; --k) {
      if (k < d.ndim) {
        int a = d.map[k];
        off += lidx[a] * acc;
        if (a == sa) stride_sa = acc;
        acc *= (k == d.ndim - 1 && d.pitch > 0) ? d.pitch : llen[a];
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < kMaxDim; ++k) {
      if (k < d.ndim) {
        int a = d.map[k];
        off += lidx[a] * acc;
        if (a == sa) stride_sa = acc;
        acc *= llen[a];
      }
    }
  }
  return d.ptr[rank] + off * esize;
}

}  // namespace dfft
