// HIP kernels for gfx950 (CDNA4): the grouped and runtime-length line passes,
// the line-pass launch entry (the compile-time-length Stockham kernels are
// templates in fft_line_impl.hpp, instantiated in the fft_lines_*.hip units),
// the exchange remap, and the reference-exact pack / unpack / all-to-all
// kernels.
//
// FFT conventions (the ones kokkos-fft exposes to the reference plans,
// KokkosFFT_Distributed_SlabPlan.hpp:978-1081): forward X[k] = sum_j x[j]
// exp(-2 pi i jk/n), backward with +i, both unscaled; scaling is applied by
// the caller-supplied factor at the final store.  Inverse transforms run the
// forward kernel between two conjugations (IDFT(x) = conj(DFT(conj(x)))).
// R2C/C2R lines are carried as full complex lines of length n (real input
// with zero imaginary part; Hermitian extension of the n/2+1 input values).
#include <hip/hip_runtime.h>

#include "fft_kernels.hpp"
#include "fft_line_impl.hpp"

#include <algorithm>
#include <cstdint>
#include <type_traits>

namespace dfft {
// ---------------------------------------------------------------------------
// Grouped line pass: the same pass of many independent transforms (lengths
// from kGroupSizes, any mix) in one launch.  Virtual blocks are numbered
// transform by transform (first_block[i] .. first_block[i+1]); a grid of at
// most a few blocks per CU walks them in a block-stride loop.  Blocks have
// kGroupThreads lanes (64 or 128 for launches of contiguous lines);
// transform i uses items[i].lines_per_block lines.  STAGE (complex64 one-line
// instances): contiguous, 16-byte-aligned lines move 16 bytes per lane through
// the line's LDS buffer.
// ---------------------------------------------------------------------------
template <typename Real, bool NT, bool DUAL, bool STAGE = false>
__global__ void __launch_bounds__(kGroupThreads) fft_group_pass(const LinePass* __restrict__ items,
                                                                const int32_t* __restrict__ first_block,
                                                                int32_t count, int32_t total_blocks) {
  extern __shared__ unsigned char smem_raw[];
  unsigned char* lds = smem_raw;
  for (int32_t vb = blockIdx.x; vb < total_blocks; vb += gridDim.x) {
    // member of virtual block vb: the last i with first_block[i] <= vb
    int32_t lo = 0, hi = count - 1;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (first_block[mid] <= vb) lo = mid;
      else hi = mid - 1;
    }
    const int32_t i = lo;
    const LinePass& p = items[i];
    const int32_t blk = vb - first_block[i];
    if constexpr (DUAL) {
      // DUAL instances hold the FP32 members taken two lines per lane (their
      // register sets are twice as large, so they are a kernel of their own)
      if constexpr (std::is_same<Real, float>::value) {
        switch (p.n) {
          case 64: fast_line2_body<Real, 64, 8, NT>(p, blk, p.lines_per_block, lds); break;
          case 96: fast_line2_body<Real, 96, 8, NT>(p, blk, p.lines_per_block, lds); break;
          case 128: fast_line2_body<Real, 128, 8, NT>(p, blk, p.lines_per_block, lds); break;
          case 160: fast_line2_body<Real, 160, 8, NT>(p, blk, p.lines_per_block, lds); break;
          case 256: fast_line2_body<Real, 256, 8, NT>(p, blk, p.lines_per_block, lds); break;
          default: break;
        }
      }
    } else {
      // complex128 elements are 16 bytes already: staging is for complex64
      constexpr bool kStage = STAGE && std::is_same<Real, float>::value;
      switch (p.n) {
        case 16: fast_line_body<Real, 16, 8, NT, false, kStage>(p, blk, p.lines_per_block, lds); break;
        case 32: fast_line_body<Real, 32, 8, NT, false, kStage>(p, blk, p.lines_per_block, lds); break;
        case 64: fast_line_body<Real, 64, 8, NT, false, kStage>(p, blk, p.lines_per_block, lds); break;
        case 96: fast_line_body<Real, 96, 8, NT, false, kStage>(p, blk, p.lines_per_block, lds); break;
        case 128: fast_line_body<Real, 128, 8, NT, false, kStage>(p, blk, p.lines_per_block, lds); break;
        case 160: fast_line_body<Real, 160, 8, NT, false, kStage>(p, blk, p.lines_per_block, lds); break;
        case 256: fast_line_body<Real, 256, 8, NT, false, kStage>(p, blk, p.lines_per_block, lds); break;
        default: break;
      }
    }
    __syncthreads();  // the next virtual block reuses the LDS lines
  }
}

// ---------------------------------------------------------------------------
// Generic runtime-length kernel (any n): lines staged in LDS, every output of
// a radix-r stage is a length-r sum with combined twiddles.  Handles lines
// that cross rank boundaries (per-element addressing).
// ---------------------------------------------------------------------------
struct GenPlan {
  int32_t nst;
  int32_t rad[32];
};

// The radix stages of B lines of length n held in bufA (line b at b*n):
// every output of a radix-r stage is an r-term sum with combined twiddles
// (tw: exp(-2 pi i e / n)).  Returns the buffer that holds the result.
template <typename Real>
__device__ cplx_t<Real>* generic_stages(cplx_t<Real>* bufA, cplx_t<Real>* bufB, int32_t n, int32_t total,
                                        const GenPlan& g, const double2* tw) {
  using C = cplx_t<Real>;
  int32_t L = 1;
  for (int s = 0; s < g.nst; ++s) {
    const int32_t r = g.rad[s];
    const int32_t Lr = L * r;
    const int32_t ts = n / Lr;
    const int32_t nr = n / r;
    for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
      const int32_t b = e / n, o = e - b * n;
      const int32_t grp = o / Lr, rm = o - grp * Lr;
      const int32_t q = rm / L, k = rm - q * L;
      const int32_t beta = grp * L + k;
      const C* src = bufA + b * n;
      C acc{Real(0), Real(0)};
      const int32_t step = k + q * L;  // exponent step in units of W_{Lr}
      int32_t ex = 0;
      for (int32_t j = 0; j < r; ++j) {
        C x = src[beta + j * nr];
        acc = cadd(acc, cmul(x, to_prec<C>(tw[ex * ts])));
        ex += step;
        if (ex >= Lr) ex -= Lr * (ex / Lr);
      }
      bufB[e] = acc;
    }
    __syncthreads();
    C* tmp = bufA;
    bufA = bufB;
    bufB = tmp;
    L = Lr;
  }
  return bufA;
}

// Input value at position pos of line `line` of pass p (per-element
// addressing): R2C lines load reals, C2R lines the Hermitian extension of
// the half spectrum (with the interior weight), inverse passes conjugate.
template <typename Real>
__device__ __forceinline__ cplx_t<Real> generic_load(const LinePass& p, int32_t line, int32_t pos) {
  using C = cplx_t<Real>;
  const int32_t n = p.n;
  int32_t idx[kMaxDim];
  decode_line<Real>(p, line, idx);
  const bool mirror = p.kind == kC2R && pos >= n / 2 + 1;
  idx[p.axis] = mirror ? n - pos : pos;
  int64_t st;
  const int64_t ies = p.kind == kR2C ? sizeof(Real) : sizeof(C);
  const char* ib = element_ptr(p.in, idx, ies, p.axis, st);
  C v;
  if (p.kind == kR2C) {
    v = C{*reinterpret_cast<const Real*>(ib), Real(0)};
  } else {
    v = *reinterpret_cast<const C*>(ib);
  }
  if (p.kind == kC2R) {
    if (mirror) v = cconj(v);
    if (pos == 0 || 2 * pos == n) {
      v.y = Real(0);
    } else if (p.herm_weight != 0.0) {
      v.x *= static_cast<Real>(p.herm_weight);
      v.y *= static_cast<Real>(p.herm_weight);
    }
  }
  return p.inverse ? cconj(v) : v;
}

// Store of the transform output v (before the inverse conjugation) at
// position pos of line `line`: scale, R2C keeps pos <= n/2 (interior weight),
// C2R stores the real part.
template <typename Real>
__device__ __forceinline__ void generic_store(const LinePass& p, int32_t line, int32_t pos, cplx_t<Real> v) {
  using C = cplx_t<Real>;
  const int32_t n = p.n;
  if (p.kind == kR2C && 2 * pos > n) return;
  if (p.inverse) v = cconj(v);
  Real sc = static_cast<Real>(p.scale);
  if (p.kind == kR2C && pos != 0 && 2 * pos != n && p.herm_weight != 0.0) sc *= static_cast<Real>(p.herm_weight);
  int32_t idx[kMaxDim];
  decode_line<Real>(p, line, idx);
  idx[p.axis] = pos;
  int64_t st;
  const int64_t oes = p.kind == kC2R ? sizeof(Real) : sizeof(C);
  char* ob = element_ptr(p.out, idx, oes, p.axis, st);
  if (p.kind == kC2R) {
    *reinterpret_cast<Real*>(ob) = v.x * sc;
  } else {
    *reinterpret_cast<C*>(ob) = C{v.x * sc, v.y * sc};
  }
}

// Lines longer than one LDS buffer (LinePass::long_n1 > 0, n = n1 * n2),
// first launch: sub-line (line, i2) holds elements i1*n2 + i2 (i1 < n1) of
// the line; its transform Y(k1) times exp(-2 pi i i2 k1 / n) goes to
// scratch[(line*n1 + k1)*n2 + i2].  Lanes walk the B sub-lines of a block
// fastest (consecutive i2: consecutive elements of the line on both sides).
template <typename Real>
__global__ void __launch_bounds__(256) fft_long_first(const LinePass p, const GenPlan g) {
  using C = cplx_t<Real>;
  extern __shared__ unsigned char smem_raw[];
  const int32_t n = p.n, n1 = p.long_n1, n2 = n / n1;
  const int B = p.lines_per_block;
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + B * n1;
  const int32_t total = B * n1;
  const int64_t nsub = static_cast<int64_t>(p.nlines) * n2;
  const int64_t sub0 = static_cast<int64_t>(blockIdx.x) * B;
  for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int32_t b = e % B, i1 = e / B;
    const int64_t sub = sub0 + b;
    C v{Real(0), Real(0)};
    if (sub < nsub) {
      const int32_t line = static_cast<int32_t>(sub / n2), i2 = static_cast<int32_t>(sub % n2);
      if (!padding_line<Real>(p, line)) v = generic_load<Real>(p, line, i1 * n2 + i2);
    }
    bufA[b * n1 + i1] = v;
  }
  __syncthreads();
  C* res = generic_stages<Real>(bufA, bufB, n1, total, g, reinterpret_cast<const double2*>(p.long_tw1));
  const double2* twn = reinterpret_cast<const double2*>(p.tw);
  C* scr = reinterpret_cast<C*>(p.long_scratch);
  for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int32_t b = e % B, k1 = e / B;
    const int64_t sub = sub0 + b;
    if (sub >= nsub) continue;
    const int32_t line = static_cast<int32_t>(sub / n2), i2 = static_cast<int32_t>(sub % n2);
    // i2 * k1 <= (n2 - 1)(n1 - 1) < n: a table index without reduction
    const C w = to_prec<C>(twn[i2 * k1]);
    scr[(static_cast<int64_t>(line) * n1 + k1) * n2 + i2] = cmul(res[b * n1 + k1], w);
  }
}

// Second launch: sub-line (line, k1) is scratch[(line*n1 + k1)*n2 + i2]
// (i2 < n2, contiguous); output k2 of its transform is element k1 + n1*k2 of
// the line.  Loads walk each sub-line, stores walk the sub-lines fastest
// (consecutive k1: consecutive elements of the output line).
template <typename Real>
__global__ void __launch_bounds__(256) fft_long_second(const LinePass p, const GenPlan g) {
  using C = cplx_t<Real>;
  extern __shared__ unsigned char smem_raw[];
  const int32_t n = p.n, n1 = p.long_n1, n2 = n / n1;
  const int B = p.lines_per_block;
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + B * n2;
  const int32_t total = B * n2;
  const int64_t nsub = static_cast<int64_t>(p.nlines) * n1;
  const int64_t sub0 = static_cast<int64_t>(blockIdx.x) * B;
  const C* scr = reinterpret_cast<const C*>(p.long_scratch);
  for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int32_t b = e / n2, i2 = e % n2;
    const int64_t sub = sub0 + b;
    bufA[e] = sub < nsub ? scr[sub * n2 + i2] : C{Real(0), Real(0)};
  }
  __syncthreads();
  C* res = generic_stages<Real>(bufA, bufB, n2, total, g, reinterpret_cast<const double2*>(p.long_tw2));
  for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int32_t b = e % B, k2 = e / B;
    const int64_t sub = sub0 + b;
    if (sub >= nsub) continue;
    const int32_t line = static_cast<int32_t>(sub / n1), k1 = static_cast<int32_t>(sub % n1);
    if (padding_line<Real>(p, line)) continue;
    generic_store<Real>(p, line, k1 + n1 * k2, res[b * n2 + k2]);
  }
}

// Bluestein (chirp-z) transform of lines whose length has a prime factor
// beyond one LDS line: with c_e = exp(-pi i e / n) (e = j^2 mod 2n),
//   X_k = c_{k^2} sum_j (x_j c_{j^2}) conj(c_{(k-j)^2}),
// a cyclic convolution of length m >= 2n - 1 (a power of two) carried out
// with two m-point transforms of the (nlines, m) scratch between these
// kernels.  bluestein_pre: s1[line][j] = x_j c_{j^2} (j < n), 0 (n <= j < m);
// the loads follow the generic kernel (R2C reals, C2R Hermitian extension,
// inverse conjugation).
template <typename Real>
__global__ void __launch_bounds__(256) bluestein_pre(const LinePass p, int32_t log2m, const double2* chirp,
                                                     cplx_t<Real>* s1) {
  using C = cplx_t<Real>;
  const int64_t m = int64_t(1) << log2m;
  const int64_t total = static_cast<int64_t>(p.nlines) << log2m;
  const int64_t n2 = 2 * static_cast<int64_t>(p.n);
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t line = static_cast<int32_t>(g >> log2m);
    const int64_t j = g & (m - 1);
    C v{Real(0), Real(0)};
    if (j < p.n && !padding_line<Real>(p, line)) {
      const C x = generic_load<Real>(p, line, static_cast<int32_t>(j));
      v = cmul(x, to_prec<C>(chirp[(j * j) % n2]));
    }
    s1[g] = v;
  }
}

// s[line][k] *= filt[k]: the m-point transform of the chirp filter, divided by m
template <typename Real>
__global__ void __launch_bounds__(256) bluestein_mul(int64_t total, int32_t log2m, const double2* filt,
                                                     cplx_t<Real>* s) {
  using C = cplx_t<Real>;
  const int64_t mask = (int64_t(1) << log2m) - 1;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x)
    s[g] = cmul(s[g], to_prec<C>(filt[g & mask]));
}

// X_k = c_{k^2} s1[line][k] (k < n), stored as the generic kernel stores
template <typename Real>
__global__ void __launch_bounds__(256) bluestein_post(const LinePass p, int32_t log2m, const double2* chirp,
                                                      const cplx_t<Real>* s1) {
  using C = cplx_t<Real>;
  const int64_t total = static_cast<int64_t>(p.nlines) * p.n;
  const int64_t n2 = 2 * static_cast<int64_t>(p.n);
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t line = static_cast<int32_t>(g / p.n);
    const int64_t k = g - static_cast<int64_t>(line) * p.n;
    if (padding_line<Real>(p, line)) continue;
    const C y = s1[(static_cast<int64_t>(line) << log2m) + k];
    generic_store<Real>(p, line, static_cast<int32_t>(k), cmul(y, to_prec<C>(chirp[(k * k) % n2])));
  }
}

template <typename Real>
__global__ void __launch_bounds__(256) fft_line_generic(const LinePass p, const GenPlan g) {
  using C = cplx_t<Real>;
  extern __shared__ unsigned char smem_raw[];
  const int32_t n = p.n;
  const int B = p.lines_per_block;
  C* bufA = reinterpret_cast<C*>(smem_raw);
  C* bufB = bufA + B * n;
  const double2* tw = reinterpret_cast<const double2*>(p.tw);
  const int32_t total = B * n;
  const int64_t ies = p.kind == kR2C ? sizeof(Real) : sizeof(C);
  const int64_t oes = p.kind == kC2R ? sizeof(Real) : sizeof(C);

  for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int32_t b = e / n, pos = e - b * n;
    const int32_t line = blockIdx.x * B + b;
    C v{Real(0), Real(0)};
    if (line < p.nlines && !padding_line<Real>(p, line)) {
      int32_t idx[kMaxDim];
      decode_line<Real>(p, line, idx);
      LineIO<Real> io;
      io.kind = p.kind;
      io.n = n;
      io.inverse = p.inverse;
      io.scale = static_cast<Real>(p.scale);
      io.hw = p.herm_weight != 0.0 ? static_cast<Real>(p.herm_weight) : Real(1);
      if (p.kind == kC2R && pos >= n / 2 + 1) {
        idx[p.axis] = n - pos;
      } else {
        idx[p.axis] = pos;
      }
      int64_t st;
      io.ib = element_ptr(p.in, idx, ies, p.axis, st);
      io.is = 0;
      // LineIO::load indexes pos*is from ib; with is = 0 the element pointer is used as is,
      // except the Hermitian extension which needs the conjugate of X[n-pos].
      if (p.kind == kC2R) {
        v = *reinterpret_cast<const C*>(io.ib);
        if (pos >= n / 2 + 1) v = cconj(v);
        if (pos == 0 || 2 * pos == n) {
          v.y = Real(0);
        } else {
          v.x *= io.hw;
          v.y *= io.hw;
        }
        if (p.inverse) v = cconj(v);
      } else {
        v = io.load(0);
      }
    }
    bufA[e] = v;
  }
  __syncthreads();
  bufA = generic_stages<Real>(bufA, bufB, n, total, g, tw);
  for (int32_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int32_t b = e / n, pos = e - b * n;
    const int32_t line = blockIdx.x * B + b;
    if (line >= p.nlines || padding_line<Real>(p, line)) continue;
    if (p.kind == kR2C && 2 * pos > n) continue;
    int32_t idx[kMaxDim];
    decode_line<Real>(p, line, idx);
    idx[p.axis] = pos;
    int64_t st;
    LineIO<Real> io;
    io.kind = p.kind;
    io.n = n;
    io.inverse = p.inverse;
    io.scale = static_cast<Real>(p.scale);
    io.hw = Real(1);
    io.ob = element_ptr(p.out, idx, oes, p.axis, st);
    io.os = 0;
    // store(0, .) sees position 0, so the interior weight is applied here
    C v = bufA[e];
    if (p.kind == kR2C && pos != 0 && 2 * pos != n && p.herm_weight != 0.0) {
      v.x *= static_cast<Real>(p.herm_weight);
      v.y *= static_cast<Real>(p.herm_weight);
    }
    io.store(0, v);
  }
}

// ---------------------------------------------------------------------------
// Launchers of the generic kernel and dispatch to the fast-kernel units
// ---------------------------------------------------------------------------
namespace {

// radices of the generic kernel for length n: 4s, then prime factors
bool make_gen_plan(int n, GenPlan& g) {
  g = GenPlan{};
  int m = n;
  while (m % 4 == 0 && m > 4) { g.rad[g.nst++] = 4; m /= 4; }
  for (int f = 2; m > 1 && g.nst < 32;) {
    if (m % f == 0) {
      g.rad[g.nst++] = f;
      m /= f;
    } else {
      ++f;
      if (f * f > m) f = m;
    }
  }
  return m == 1;
}

// sub-lines per block of a long-line launch: lines of length len, twice, in
// at most 64 KB of LDS (at least one line, up to 160 KB)
template <typename Real>
int long_block(int len) {
  const size_t line = static_cast<size_t>(2) * len * sizeof(cplx_t<Real>);
  int B = 64;
  while (B > 1 && static_cast<size_t>(B) * line > 64 * 1024) B /= 2;
  return B;
}

template <typename Real>
hipError_t launch_long(LinePass p, hipStream_t s) {
  using C = cplx_t<Real>;
  const int n1 = p.long_n1;
  if (n1 <= 0 || p.n % n1 != 0 || !p.long_scratch || !p.long_tw1 || !p.long_tw2) return hipErrorInvalidValue;
  const int n2 = p.n / n1;
  GenPlan g1, g2;
  if (!make_gen_plan(n1, g1) || !make_gen_plan(n2, g2)) return hipErrorInvalidValue;
  static std::atomic<size_t> lds1{64 * 1024}, lds2{64 * 1024};
  {
    LinePass q = p;
    q.lines_per_block = long_block<Real>(n1);
    const size_t smem = static_cast<size_t>(2) * q.lines_per_block * n1 * sizeof(C);
    if (smem > kGenericLdsMax) return hipErrorInvalidValue;
    const hipError_t e = raise_lds_limit(lds1, reinterpret_cast<const void*>(&fft_long_first<Real>), smem);
    if (e != hipSuccess) return e;
    const int64_t grid = (static_cast<int64_t>(p.nlines) * n2 + q.lines_per_block - 1) / q.lines_per_block;
    if (grid >= (int64_t(1) << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((fft_long_first<Real>), dim3(static_cast<unsigned>(grid)), dim3(256), smem, s, q, g1);
    const hipError_t e1 = hipGetLastError();
    if (e1 != hipSuccess) return e1;
  }
  LinePass q = p;
  q.lines_per_block = long_block<Real>(n2);
  const size_t smem = static_cast<size_t>(2) * q.lines_per_block * n2 * sizeof(C);
  if (smem > kGenericLdsMax) return hipErrorInvalidValue;
  const hipError_t e = raise_lds_limit(lds2, reinterpret_cast<const void*>(&fft_long_second<Real>), smem);
  if (e != hipSuccess) return e;
  const int64_t grid = (static_cast<int64_t>(p.nlines) * n1 + q.lines_per_block - 1) / q.lines_per_block;
  if (grid >= (int64_t(1) << 31)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((fft_long_second<Real>), dim3(static_cast<unsigned>(grid)), dim3(256), smem, s, q, g2);
  return hipGetLastError();
}

unsigned elementwise_grid(int64_t total) {
  const int64_t blocks = (total + 255) / 256;
  return static_cast<unsigned>(std::min<int64_t>(blocks, 65536));
}

template <typename Real>
hipError_t launch_generic(LinePass p, hipStream_t s) {
  using C = cplx_t<Real>;
  GenPlan g{};
  if (!make_gen_plan(p.n, g)) return hipErrorInvalidValue;
  int B = p.n >= 1024 ? 1 : 1024 / p.n;
  while (B > 1 && static_cast<size_t>(2) * B * p.n * sizeof(C) > 64 * 1024) B /= 2;
  const size_t smem = static_cast<size_t>(2) * B * p.n * sizeof(C);
  if (smem > kGenericLdsMax) return hipErrorInvalidValue;
  static std::atomic<size_t> lds_limit{64 * 1024};
  {
    const hipError_t e = raise_lds_limit(lds_limit, reinterpret_cast<const void*>(&fft_line_generic<Real>), smem);
    if (e != hipSuccess) return e;
  }
  p.lines_per_block = B;
  const unsigned grid = static_cast<unsigned>((p.nlines + B - 1) / B);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((fft_line_generic<Real>), dim3(grid), dim3(256), smem, s, p, g);
  return hipGetLastError();
}

template <typename Real>
hipError_t dispatch(const LinePass& p, hipStream_t s) {
  const bool fp64 = std::is_same<Real, double>::value;
  if (p.n <= 160) return fp64 ? dispatch_fast_f64_short(p, s) : dispatch_fast_f32_short(p, s);
  if (p.n >= 4096) return fp64 ? dispatch_fast_f64_xl(p, s) : dispatch_fast_f32_xl(p, s);
  return fp64 ? dispatch_fast_f64_long(p, s) : dispatch_fast_f32_long(p, s);
}

}  // namespace

namespace {
// lines per block: as many as the block's lanes hold, capped so the lines
// plus the spare slot fit `cap` bytes of dynamic LDS (at most 64 KB)
template <typename Real, int N>
void group_shape_of(size_t cap, int lanes, int* B, size_t* lds) {
  using Cfg = FastCfg<Real, N, 8>;
  static_assert(Cfg::T <= 64, "a grouped line fits the smallest block");
  const size_t line_bytes = static_cast<size_t>(Cfg::LDSN) * sizeof(cplx_t<Real>);
  if (cap > kGroupLds) cap = kGroupLds;
  int b = lanes / Cfg::T;
  while (b > 1 && static_cast<size_t>(b + 1) * line_bytes > cap) --b;
  *B = b;
  *lds = static_cast<size_t>(b + 1) * line_bytes;
}
template <typename Real>
bool group_shape(int n, size_t cap, int lanes, int* B, size_t* lds) {
  switch (n) {
    case 16: group_shape_of<Real, 16>(cap, lanes, B, lds); return true;
    case 32: group_shape_of<Real, 32>(cap, lanes, B, lds); return true;
    case 64: group_shape_of<Real, 64>(cap, lanes, B, lds); return true;
    case 96: group_shape_of<Real, 96>(cap, lanes, B, lds); return true;
    case 128: group_shape_of<Real, 128>(cap, lanes, B, lds); return true;
    case 160: group_shape_of<Real, 160>(cap, lanes, B, lds); return true;
    case 256: group_shape_of<Real, 256>(cap, lanes, B, lds); return true;
    default: return false;
  }
}
}  // namespace

namespace {
// two lines per lane: pairs = lanes / T, lines = 2 * pairs, LDS for the
// lines and the two spare slots (real / imaginary halves of each line)
template <int N> void group_dual_shape_of(size_t cap, int* B, size_t* lds) {
  using Cfg = FastCfg<float, N, 8>;
  const size_t line_bytes = dual_line_bytes<float, N, 8>();
  if (cap > kGroupLds) cap = kGroupLds;
  int pairs = kGroupThreads / Cfg::T;
  while (pairs > 1 && static_cast<size_t>(2 * pairs + 2) * line_bytes > cap) --pairs;
  *B = 2 * pairs;
  *lds = static_cast<size_t>(2 * pairs + 2) * line_bytes;
}
}  // namespace

bool group_dual_shape(int n, size_t lds_cap, int* lines_per_block, size_t* lds_bytes) {
  switch (n) {
    case 64: group_dual_shape_of<64>(lds_cap, lines_per_block, lds_bytes); return true;
    case 96: group_dual_shape_of<96>(lds_cap, lines_per_block, lds_bytes); return true;
    case 128: group_dual_shape_of<128>(lds_cap, lines_per_block, lds_bytes); return true;
    case 160: group_dual_shape_of<160>(lds_cap, lines_per_block, lds_bytes); return true;
    case 256: group_dual_shape_of<256>(lds_cap, lines_per_block, lds_bytes); return true;
    default: return false;
  }
}

bool dual_lines_apply(const LinePass& p, int lines_per_block) { return dual_lines_apply_impl(p, lines_per_block); }

bool group_line_shape(int n, bool fp64, size_t lds_cap, int lanes, int* lines_per_block, size_t* lds_bytes) {
  if (lanes != 64 && lanes != 128 && lanes != kGroupThreads) return false;
  return fp64 ? group_shape<double>(n, lds_cap, lanes, lines_per_block, lds_bytes)
              : group_shape<float>(n, lds_cap, lanes, lines_per_block, lds_bytes);
}

int group_grid(size_t lds_bytes, bool fp64, bool dual, int lanes) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const void* f = fp64   ? reinterpret_cast<const void*>(&fft_group_pass<double, false, false>)
                  : dual ? reinterpret_cast<const void*>(&fft_group_pass<float, false, true>)
                         : reinterpret_cast<const void*>(&fft_group_pass<float, false, false, true>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, dual ? kGroupThreads : lanes, lds_bytes) != hipSuccess)
    return 0;
  return cus * (per_cu > 0 ? per_cu : 1);
}

hipError_t launch_group_pass(const LinePass* items, const int32_t* first_block, int count, int total_blocks,
                             int grid, size_t lds_bytes, bool fp64, bool nontemporal, bool dual, int lanes,
                             bool stage, hipStream_t s) {
  if (total_blocks <= 0 || count <= 0) return hipSuccess;
  if (grid <= 0 || lds_bytes > kGroupLds || (dual && fp64)) return hipErrorInvalidConfiguration;
  if (lanes != 64 && lanes != 128 && lanes != kGroupThreads) return hipErrorInvalidConfiguration;
  if (dual && lanes != kGroupThreads) return hipErrorInvalidConfiguration;
  const dim3 g(static_cast<unsigned>(grid < total_blocks ? grid : total_blocks));
  const dim3 b(static_cast<unsigned>(lanes));
  if (fp64) {
    if (nontemporal)
      hipLaunchKernelGGL((fft_group_pass<double, true, false>), g, b, lds_bytes, s, items, first_block, count,
                         total_blocks);
    else
      hipLaunchKernelGGL((fft_group_pass<double, false, false>), g, b, lds_bytes, s, items, first_block, count,
                         total_blocks);
  } else if (dual) {
    if (nontemporal)
      hipLaunchKernelGGL((fft_group_pass<float, true, true>), g, b, lds_bytes, s, items, first_block, count,
                         total_blocks);
    else
      hipLaunchKernelGGL((fft_group_pass<float, false, true>), g, b, lds_bytes, s, items, first_block, count,
                         total_blocks);
  } else if (stage) {
    if (nontemporal)
      hipLaunchKernelGGL((fft_group_pass<float, true, false, true>), g, b, lds_bytes, s, items, first_block, count,
                         total_blocks);
    else
      hipLaunchKernelGGL((fft_group_pass<float, false, false, true>), g, b, lds_bytes, s, items, first_block,
                         count, total_blocks);
  } else {
    if (nontemporal)
      hipLaunchKernelGGL((fft_group_pass<float, true, false>), g, b, lds_bytes, s, items, first_block, count,
                         total_blocks);
    else
      hipLaunchKernelGGL((fft_group_pass<float, false, false>), g, b, lds_bytes, s, items, first_block, count,
                         total_blocks);
  }
  return hipGetLastError();
}

bool has_fast_line_kernel(int n) {
  switch (n) {
    case 2: case 4: case 8: case 16: case 32: case 64: case 96: case 128: case 160: case 256:
    case 512: case 1024: case 2048: case 4096: case 8192:
      return true;
    default:
      return false;
  }
}

bool line_length_single_pass(int n, bool fp64) {
  if (n < 1) return false;
  if (has_fast_line_kernel(n)) return true;
  const size_t csize = fp64 ? sizeof(double2) : sizeof(float2);
  return static_cast<size_t>(2) * static_cast<size_t>(n) * csize <= kGenericLdsMax;
}

int long_line_split(int n, bool fp64) {
  if (n < 4 || line_length_single_pass(n, fp64)) return 0;
  int best = 0;
  for (int d = 2; static_cast<int64_t>(d) * d <= n; ++d)
    if (n % d == 0 && line_length_single_pass(d, fp64) && line_length_single_pass(n / d, fp64)) best = d;
  return best;
}

bool line_length_supported(int n, bool fp64) {
  return line_length_single_pass(n, fp64) || long_line_split(n, fp64) > 0;
}

// Cost model (ms per 16.8M complex elements, forward + backward, measured on
// (n, W) arrays with prime n from 11 to 4099): the generic kernel costs
// 0.5 + 0.031 * (sum of its radices) -- its r-term sums make a prime p cost
// ~p operations per element -- in both precisions; Bluestein costs
// c * m / n + 0.3 with c = 1.15 (FP64) / 0.7 (FP32) for m <= 2048, 2.7 / 1.5
// for m = 4096 and 8192 (one-line kernels of 256 and 512 lanes) and 10.9 /
// 12 beyond (two launches of the generic kernel).  Measured pairs (generic /
// Bluestein, FP64): 61: 2.05 / 2.21, 127: 3.84 / 2.64, 1009: 34.8 / 2.6,
// 2053: 94 / 5.7, 4099: 376 / 21.7 ms; FP32 31: 1.09 / 1.39, 61: 1.87 / 1.41,
// 2053: 79 / 3.4 ms.  Bluestein is taken where it models at least 20 %
// cheaper than every other kernel that can run the length.
bool bluestein_preferred(int n, bool fp64) {
  if (n < 2 || has_fast_line_kernel(n)) return false;
  if (!line_length_supported(n, fp64)) return true;
  auto generic = [](int len) {
    GenPlan g;
    if (!make_gen_plan(len, g)) return 1e30;
    int sum = 0;
    for (int s = 0; s < g.nst; ++s) sum += g.rad[s];
    return 0.5 + 0.031 * sum;
  };
  auto pass_cost = [&](int len) { return has_fast_line_kernel(len) ? 0.5 : generic(len); };
  double other;
  const int n1 = long_line_split(n, fp64);
  if (n1 > 0) other = pass_cost(n1) + pass_cost(n / n1) + 0.5;  // two launches and their twiddle
  else other = generic(n);
  int64_t m = 1;
  while (m < 2 * static_cast<int64_t>(n) - 1) m *= 2;
  const double c = m <= 2048 ? (fp64 ? 1.15 : 0.7) : m <= 8192 ? (fp64 ? 2.7 : 1.5) : (fp64 ? 10.9 : 12.0);
  const double blue = c * static_cast<double>(m) / n + 0.3;
  return blue < 0.8 * other;
}

bool products_pass_supported(int n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }

hipError_t launch_products_pass(const LinePass& p, hipStream_t s) {
  if (p.nlines <= 0) return hipSuccess;
  if (!products_pass_supported(p.n) || p.in.ntopo[p.axis] != 1 || p.out.ntopo[p.axis] != 1)
    return hipErrorNotSupported;
  return dispatch_products_f64(p, s);
}

hipError_t launch_bluestein_pre(const LinePass& p, int log2m, const void* chirp, void* s1, bool fp64,
                                hipStream_t s) {
  if (p.nlines <= 0) return hipSuccess;
  const unsigned grid = elementwise_grid(static_cast<int64_t>(p.nlines) << log2m);
  const double2* c = static_cast<const double2*>(chirp);
  if (fp64) hipLaunchKernelGGL((bluestein_pre<double>), dim3(grid), dim3(256), 0, s, p, log2m, c, static_cast<double2*>(s1));
  else hipLaunchKernelGGL((bluestein_pre<float>), dim3(grid), dim3(256), 0, s, p, log2m, c, static_cast<float2*>(s1));
  return hipGetLastError();
}

hipError_t launch_bluestein_mul(int64_t total, int log2m, const void* filt, void* buf, bool fp64, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  const unsigned grid = elementwise_grid(total);
  const double2* f = static_cast<const double2*>(filt);
  if (fp64) hipLaunchKernelGGL((bluestein_mul<double>), dim3(grid), dim3(256), 0, s, total, log2m, f, static_cast<double2*>(buf));
  else hipLaunchKernelGGL((bluestein_mul<float>), dim3(grid), dim3(256), 0, s, total, log2m, f, static_cast<float2*>(buf));
  return hipGetLastError();
}

hipError_t launch_bluestein_post(const LinePass& p, int log2m, const void* chirp, const void* s1, bool fp64,
                                 hipStream_t s) {
  if (p.nlines <= 0) return hipSuccess;
  const unsigned grid = elementwise_grid(static_cast<int64_t>(p.nlines) * p.n);
  const double2* c = static_cast<const double2*>(chirp);
  if (fp64)
    hipLaunchKernelGGL((bluestein_post<double>), dim3(grid), dim3(256), 0, s, p, log2m, c, static_cast<const double2*>(s1));
  else
    hipLaunchKernelGGL((bluestein_post<float>), dim3(grid), dim3(256), 0, s, p, log2m, c, static_cast<const float2*>(s1));
  return hipGetLastError();
}

hipError_t launch_line_pass(const LinePass& p, bool fp64, hipStream_t s) {
  if (p.nlines <= 0) return hipSuccess;
  if (p.long_n1 > 0) return fp64 ? launch_long<double>(p, s) : launch_long<float>(p, s);
  // the compile-time kernel needs the line axis undistributed on both sides
  const bool fast = has_fast_line_kernel(p.n) && p.in.ntopo[p.axis] == 1 && p.out.ntopo[p.axis] == 1;
  if (fast) return fp64 ? dispatch<double>(p, s) : dispatch<float>(p, s);
  return fp64 ? launch_generic<double>(p, s) : launch_generic<float>(p, s);
}

// ---------------------------------------------------------------------------
// Remap (one-GPU exchange / local transpose)
// ---------------------------------------------------------------------------
template <typename E>
__global__ void __launch_bounds__(256) remap_kernel(const RemapArgs a) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < a.total; g += stride) {
    int32_t idx[kMaxDim] = {0, 0, 0, 0};
    uint32_t rem = static_cast<uint32_t>(g);
#pragma unroll
    for (int k = kMaxDim - 1; k >= 0; --k) {
      if (k < a.dst.ndim) {
        uint32_t q = a.iter_div[k].div(rem);
        idx[a.iter_axis[k]] = static_cast<int32_t>(rem - q * static_cast<uint32_t>(a.iter_ext[k]));
        rem = q;
      }
    }
    int64_t st;
    const E* sp = reinterpret_cast<const E*>(element_ptr(a.src, idx, sizeof(E), 0, st));
    E* dp = reinterpret_cast<E*>(element_ptr(a.dst, idx, sizeof(E), 0, st));
    *dp = *sp;
  }
}

hipError_t launch_remap(const RemapArgs& a, int esize, hipStream_t s) {
  if (a.total <= 0) return hipSuccess;
  int64_t blocks = (a.total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  switch (esize) {
    case 4: hipLaunchKernelGGL(remap_kernel<uint32_t>, dim3(blocks), dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL(remap_kernel<uint2>, dim3(blocks), dim3(256), 0, s, a); break;
    case 16: hipLaunchKernelGGL(remap_kernel<uint4>, dim3(blocks), dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pack / unpack exactly as KokkosFFT_Distributed_PackUnpack.hpp:217-233 and
// :385-417 (LayoutRight buffer (P, e_0..e_{D-1}), logical axis order).
// ---------------------------------------------------------------------------
template <typename E, bool UNPACK>
__global__ void __launch_bounds__(256) pack_kernel(const PackArgs a) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < a.total; g += stride) {
    // j[0] = p, j[1 + i] = buffer index along logical axis i; a LayoutRight
    // buffer is (P, e_0..e_{D-1}) row-major, a LayoutLeft one (e_0..e_{D-1}, P)
    // column-major (buf_ext in that order)
    int64_t j[kMaxDim + 1];
    int64_t rem = g;
    if (!a.layout_left) {
      for (int k = a.ndim; k >= 0; --k) {
        j[k] = rem % a.buf_ext[k];
        rem /= a.buf_ext[k];
      }
    } else {
      for (int k = 0; k <= a.ndim; ++k) {
        const int64_t v = rem % a.buf_ext[k];
        rem /= a.buf_ext[k];
        if (k < a.ndim) j[k + 1] = v;
        else j[0] = v;
      }
    }
    const int64_t p = j[0];
    const int64_t base = a.merged_size / a.nprocs, r = a.merged_size % a.nprocs;
    const int64_t len = base + (p < r ? 1 : 0);
    const int64_t start = p * base + (p < r ? p : r);
    int64_t l[kMaxDim];
    bool ok = true;
    for (int i = 0; i < a.ndim; ++i) {
      if (i == a.axis) {
        if (j[i + 1] >= len) ok = false;
        l[i] = j[i + 1] + start;
      } else {
        l[i] = j[i + 1];
      }
    }
    // storage offset in the local array (storage axis k holds logical axis
    // map[k]; row-major for LayoutRight, column-major for LayoutLeft)
    int64_t off = 0;
    for (int i = 0; i < a.ndim; ++i) {
      const int k = a.layout_left ? a.ndim - 1 - i : i;
      const int64_t v = l[a.map[k]];
      if (v >= a.local_ext[k]) ok = false;
      off = off * a.local_ext[k] + v;
    }
    if constexpr (UNPACK) {
      if (ok) reinterpret_cast<E*>(a.dst)[off] = reinterpret_cast<const E*>(a.src)[g];
    } else {
      E v{};
      if (ok) v = reinterpret_cast<const E*>(a.src)[off];
      reinterpret_cast<E*>(a.dst)[g] = v;
    }
  }
}

hipError_t launch_pack(const PackArgs& a, int esize, bool unpack, hipStream_t s) {
  if (a.total <= 0) return hipSuccess;
  int64_t blocks = (a.total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
#define DFFT_PACK_CASE(T)                                                                   \
  if (unpack) hipLaunchKernelGGL((pack_kernel<T, true>), dim3(blocks), dim3(256), 0, s, a); \
  else hipLaunchKernelGGL((pack_kernel<T, false>), dim3(blocks), dim3(256), 0, s, a);
  switch (esize) {
    case 4: DFFT_PACK_CASE(uint32_t) break;
    case 8: DFFT_PACK_CASE(uint2) break;
    case 16: DFFT_PACK_CASE(uint4) break;
    default: return hipErrorInvalidValue;
  }
#undef DFFT_PACK_CASE
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) all2all_kernel(const All2AllArgs a, int64_t words_per_block) {
  const int64_t total = static_cast<int64_t>(a.nprocs) * a.nprocs * words_per_block;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t w = g % words_per_block;
    const int64_t pq = g / words_per_block;
    const int q = static_cast<int>(pq / a.nprocs);  // receiving rank
    const int p = static_cast<int>(pq % a.nprocs);  // sending rank
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.send[p] + q * a.block_bytes);
    uint32_t* dst = reinterpret_cast<uint32_t*>(a.recv[q] + p * a.block_bytes);
    dst[w] = src[w];
  }
}

hipError_t launch_all2all(const All2AllArgs& a, hipStream_t s) {
  if (a.block_bytes % 4 != 0) return hipErrorInvalidValue;
  const int64_t words = a.block_bytes / 4;
  const int64_t total = static_cast<int64_t>(a.nprocs) * a.nprocs * words;
  if (total == 0) return hipSuccess;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(all2all_kernel, dim3(blocks), dim3(256), 0, s, a, words);
  return hipGetLastError();
}

}  // namespace dfft
