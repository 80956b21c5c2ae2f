// Templates of the compile-time-length line-pass kernels and their
// launchers, shared by the translation units that instantiate them (one per
// precision and range of lengths, compiled in parallel).
//
// FFT conventions (the ones kokkos-fft exposes to the reference plans,
// KokkosFFT_Distributed_SlabPlan.hpp:978-1081): forward X[k] = sum_j x[j]
// exp(-2 pi i jk/n), backward with +i, both unscaled; scaling is applied by
// the caller-supplied factor at the final store.  Inverse transforms run the
// forward kernel between two conjugations (IDFT(x) = conj(DFT(conj(x)))).
#pragma once

#include <hip/hip_runtime.h>

#include "fft_kernels.hpp"

#include <atomic>
#include <cstdint>
#include <mutex>
#include <type_traits>

namespace dfft {

// ---------------------------------------------------------------------------
// Radix plans of the compile-time kernels.
// ---------------------------------------------------------------------------
struct FPlan {
  int R = 1;    // complex values each thread keeps in registers
  int nst = 0;  // number of stages
  int rad[12] = {};
  bool ok = true;
};

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }

// Power-of-two stages first, then 3s and 5s.  R (values per thread) is the
// least common multiple of the radices, doubled while below `rmin` and still
// a divisor of n.  With an odd factor the power-of-two part uses radix 4 (not
// 8) so R stays small: 160 = 4*4*2*5 holds 20 values per thread, not 40.
constexpr FPlan make_plan(int n, int rmin = 1) {
  FPlan p;
  if (n <= 1) return p;
  int m = n;
  int odd = n;
  while (odd % 2 == 0) odd /= 2;
  if (odd == 1) {
    while (m % 8 == 0) { p.rad[p.nst++] = 8; m /= 8; }
  }
  while (m % 4 == 0) { p.rad[p.nst++] = 4; m /= 4; }
  if (m % 2 == 0) { p.rad[p.nst++] = 2; m /= 2; }
  while (m % 3 == 0) { p.rad[p.nst++] = 3; m /= 3; }
  while (m % 5 == 0) { p.rad[p.nst++] = 5; m /= 5; }
  p.ok = (m == 1);
  int R = 1;
  for (int s = 0; s < p.nst; ++s) R = R / cgcd(R, p.rad[s]) * p.rad[s];
  while (R < rmin && n % (2 * R) == 0) R *= 2;
  p.R = R;
  return p;
}

constexpr int span_before(int n, int s) {
  FPlan p = make_plan(n);
  int L = 1;
  for (int i = 0; i < s; ++i) L *= p.rad[i];
  return L;
}

// ---------------------------------------------------------------------------
// Forward DFT butterflies on registers (exponent -2 pi i jk/r).
// ---------------------------------------------------------------------------
template <typename C> __device__ __forceinline__ void bfly2(C* a) {
  C t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}

template <typename C> __device__ __forceinline__ void bfly4(C* a) {
  C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
  C t2 = cadd(a[1], a[3]), t3 = cmul_mi(csub(a[1], a[3]));
  a[0] = cadd(t0, t2);
  a[2] = csub(t0, t2);
  a[1] = cadd(t1, t3);
  a[3] = csub(t1, t3);
}

template <typename C> __device__ __forceinline__ void bfly8(C* a) {
  using R = decltype(a[0].x);
  const R h = static_cast<R>(0.70710678118654752440);  // 1/sqrt(2)
  C e[4] = {a[0], a[2], a[4], a[6]};
  C o[4] = {a[1], a[3], a[5], a[7]};
  bfly4(e);
  bfly4(o);
  // o[k] *= exp(-2 pi i k / 8)
  C o1 = C{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
  C o2 = cmul_mi(o[2]);
  C o3 = C{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
  a[0] = cadd(e[0], o[0]);
  a[4] = csub(e[0], o[0]);
  a[1] = cadd(e[1], o1);
  a[5] = csub(e[1], o1);
  a[2] = cadd(e[2], o2);
  a[6] = csub(e[2], o2);
  a[3] = cadd(e[3], o3);
  a[7] = csub(e[3], o3);
}

template <typename C> __device__ __forceinline__ void bfly3(C* a) {
  using R = decltype(a[0].x);
  const R s3 = static_cast<R>(0.86602540378443864676);  // sin(2 pi / 3)
  C s = cadd(a[1], a[2]);
  C d = csub(a[1], a[2]);
  C m = C{a[0].x - static_cast<R>(0.5) * s.x, a[0].y - static_cast<R>(0.5) * s.y};
  C t = cmul_mi(C{s3 * d.x, s3 * d.y});
  a[0] = cadd(a[0], s);
  a[1] = cadd(m, t);
  a[2] = csub(m, t);
}

template <typename C> __device__ __forceinline__ void bfly5(C* a) {
  using R = decltype(a[0].x);
  const R c1 = static_cast<R>(0.30901699437494742410);   // cos(2 pi/5)
  const R c2 = static_cast<R>(-0.80901699437494742410);  // cos(4 pi/5)
  const R s1 = static_cast<R>(0.95105651629515357212);   // sin(2 pi/5)
  const R s2 = static_cast<R>(0.58778525229247312917);   // sin(4 pi/5)
  C b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
  C d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
  C A1 = C{a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y};
  C A2 = C{a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y};
  C X1 = cmul_mi(C{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y});
  C X2 = cmul_mi(C{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y});
  a[0] = cadd(a[0], cadd(b1, b2));
  a[1] = cadd(A1, X1);
  a[4] = csub(A1, X1);
  a[2] = cadd(A2, X2);
  a[3] = csub(A2, X2);
}

template <int r, typename C> __device__ __forceinline__ void bfly(C* a) {
  if constexpr (r == 2) bfly2(a);
  else if constexpr (r == 3) bfly3(a);
  else if constexpr (r == 4) bfly4(a);
  else if constexpr (r == 5) bfly5(a);
  else if constexpr (r == 8) bfly8(a);
}

// ---------------------------------------------------------------------------
// Element access of a line.
// ---------------------------------------------------------------------------
// Global element access: nontemporal (streaming) or cached, chosen at
// compile time (a runtime choice between the two forms of one access is
// folded into the cached form by the compiler).  The pointer is taken as a
// global-memory one: line addresses read from a table in device memory (the
// grouped kernel) are otherwise generic and compile to flat instructions,
// which also count against the LDS wait counter.
template <typename T> using global_ptr = __attribute__((address_space(1))) T*;
template <bool NT, typename T> __device__ __forceinline__ T gload(const T* p) {
  const global_ptr<const T> g = (global_ptr<const T>)p;
  if constexpr (NT) return __builtin_nontemporal_load(g);
  else return *g;
}
template <bool NT, typename T> __device__ __forceinline__ void gstore(T* p, T v) {
  const global_ptr<T> g = (global_ptr<T>)p;
  if constexpr (NT) __builtin_nontemporal_store(v, g);
  else *g = v;
}
template <bool NT, typename Real> __device__ __forceinline__ cplx_t<Real> gload_c(const cplx_t<Real>* p) {
  typedef Real V2 __attribute__((ext_vector_type(2)));
  const V2 t = gload<NT>(reinterpret_cast<const V2*>(p));
  return cplx_t<Real>{t.x, t.y};
}
template <bool NT, typename Real> __device__ __forceinline__ void gstore_c(cplx_t<Real>* p, cplx_t<Real> v) {
  typedef Real V2 __attribute__((ext_vector_type(2)));
  V2 t;
  t.x = v.x;
  t.y = v.y;
  gstore<NT>(reinterpret_cast<V2*>(p), t);
}

template <typename Real, bool NT = false> struct LineIO {
  using C = cplx_t<Real>;
  const char* ib;
  char* ob;
  int64_t is, os;
  int32_t kind, n, inverse;
  Real scale;
  Real hw;  // interior half-spectrum weight (1: none)

  __device__ __forceinline__ C load(int32_t pos) const {
    if (kind == kC2C) return load_kind<kC2C>(pos);
    if (kind == kR2C) return load_kind<kR2C>(pos);
    return load_kind<kC2R>(pos);
  }
  // the same for a kind known at compile time (a loop over positions then
  // has no per-element branch on the kind)
  template <int KIND> __device__ __forceinline__ C load_kind(int32_t pos) const {
    C v;
    if constexpr (KIND == kC2C) {
      v = gload_c<NT, Real>(reinterpret_cast<const C*>(ib) + pos * is);
    } else if constexpr (KIND == kR2C) {
      v = C{gload<NT>(reinterpret_cast<const Real*>(ib) + pos * is), Real(0)};
    } else {
      // the address is selected, not the load: loads under a branch each
      // would be waited for one by one
      const int32_t nh = n / 2 + 1;
      const bool upper = pos >= nh;
      v = gload_c<NT, Real>(reinterpret_cast<const C*>(ib) + (upper ? n - pos : pos) * is);
      if (upper) v = cconj(v);
      if (pos == 0 || 2 * pos == n) {
        v.y = Real(0);
      } else {
        v.x *= hw;
        v.y *= hw;
      }
    }
    return inverse ? cconj(v) : v;
  }
  __device__ __forceinline__ void store(int32_t pos, C v) const {
    if (inverse) v = cconj(v);
    v.x *= scale;
    v.y *= scale;
    if (kind == kC2C) {
      gstore_c<NT, Real>(reinterpret_cast<C*>(ob) + pos * os, v);
    } else if (kind == kR2C) {
      if (pos != 0 && 2 * pos != n) {
        v.x *= hw;
        v.y *= hw;
      }
      if (2 * pos <= n) gstore_c<NT, Real>(reinterpret_cast<C*>(ob) + pos * os, v);
    } else {
      gstore<NT>(reinterpret_cast<Real*>(ob) + pos * os, v.x);
    }
  }
};

// Two real lines a, b per complex line.  R2C: z = a + i b, and after the
// forward FFT Z the spectra are A_k = (Z_k + conj Z_{n-k}) / 2 and
// B_k = (Z_k - conj Z_{n-k}) / (2i).  C2R: Z = A + i B over the Hermitian
// extensions of both half spectra; the inverse transform of Z is a + i b.
template <typename Real, bool NT = false> struct PairIO {
  using C = cplx_t<Real>;
  const char* ib[2];  // products passes: ib[0] = the velocity line this slot stages
  char* ob[2];
  int64_t is[2], os[2];  // per line: the two lines may lie in different rank blocks
  int32_t kind, n;
  Real scale;
  Real hw;     // interior half-spectrum weight (1: none)
  bool has_b;  // the pair's second line exists

  // C2R input at position pos (C2R lines are conjugated: inverse = conj o
  // FFT o conj) from the loaded half-spectrum entries A = X_a[src], Bv =
  // X_b[src], src = pos below n/2+1 (lo) and n - pos above
  __device__ __forceinline__ C c2r_value(int32_t pos, bool lo, C A, C Bv) const {
    if (!lo) {
      A = cconj(A);
      Bv = cconj(Bv);
    }
    if (pos == 0 || 2 * pos == n) {
      A.y = Real(0);
      Bv.y = Real(0);
    } else {
      A.x *= hw;
      A.y *= hw;
      Bv.x *= hw;
      Bv.y *= hw;
    }
    // conj(A + i B)
    return C{A.x - Bv.y, -(A.y + Bv.x)};
  }
  // R2C: spectra of both lines at pos <= n/2 from Z_pos and Z_{n-pos}
  __device__ __forceinline__ void r2c_values(int32_t pos, C zk, C zm, C& A, C& Bv) const {
    const Real h = Real(0.5) * scale * ((pos != 0 && 2 * pos != n) ? hw : Real(1));
    A = C{(zk.x + zm.x) * h, (zk.y - zm.y) * h};
    Bv = C{(zk.y + zm.y) * h, (zm.x - zk.x) * h};
  }
};

template <typename Real>
__device__ __forceinline__ void decode_line(const LinePass& p, int32_t line, int32_t* idx) {
#pragma unroll
  for (int a = 0; a < kMaxDim; ++a) idx[a] = 0;
  uint32_t rem = static_cast<uint32_t>(line + p.line_begin);
#pragma unroll
  for (int j = kMaxDim - 2; j >= 0; --j) {
    if (j < p.nbatch) {
      uint32_t q = p.batch_div[j].div(rem);
      idx[p.batch_axis[j]] = static_cast<int32_t>(rem - q * static_cast<uint32_t>(p.batch_ext[j]));
      rem = q;
    }
  }
}

// true if line `line` of a pass enumerated over a padded extent is padding
template <typename Real>
__device__ __forceinline__ bool padding_line(const LinePass& p, int32_t line) {
  if (p.pad_real <= 0) return false;
  int32_t idx[kMaxDim];
  decode_line<Real>(p, line, idx);
  return idx[p.pad_axis] >= p.pad_real;
}

template <typename Real, bool NT = false>
__device__ __forceinline__ LineIO<Real, NT> make_line_io(const LinePass& p, int32_t line) {
  using C = cplx_t<Real>;
  int32_t idx[kMaxDim];
  decode_line<Real>(p, line, idx);
  LineIO<Real, NT> io;
  const int64_t ies = p.kind == kR2C ? sizeof(Real) : sizeof(C);
  const int64_t oes = p.kind == kC2R ? sizeof(Real) : sizeof(C);
  io.ib = element_ptr(p.in, idx, ies, p.axis, io.is);
  io.ob = element_ptr(p.out, idx, oes, p.axis, io.os);
  io.kind = p.kind;
  io.n = p.n;
  io.inverse = p.inverse;
  io.scale = static_cast<Real>(p.scale);
  io.hw = p.herm_weight != 0.0 ? static_cast<Real>(p.herm_weight) : Real(1);
  return io;
}

// factors of product field f (0..5) of the divergence-form products
__device__ __forceinline__ int32_t prod_first(int32_t f) { return f < 3 ? 0 : (f < 5 ? 1 : 2); }
__device__ __forceinline__ int32_t prod_second(int32_t f) { return f < 3 ? f : (f < 5 ? f - 2 : 2); }

// Products pass: line `idx` is product field f = idx[prod_axis] of one
// position; the velocity field (f / 2) of that position, which the slot of
// the pair (f, f + 1) stages for the position's three slots.  Selects over the
// unrolled axes keep idx in registers.
__device__ __forceinline__ const char* product_source(const LinePass& p, const int32_t* idx, int64_t ies,
                                                      int64_t& is) {
  int32_t f = 0;
#pragma unroll
  for (int a = 0; a < kMaxDim; ++a)
    if (a == p.prod_axis) f = idx[a];
  int32_t j[kMaxDim];
#pragma unroll
  for (int a = 0; a < kMaxDim; ++a) j[a] = a == p.prod_axis ? f / 2 : idx[a];
  return element_ptr(p.in, j, ies, p.axis, is);
}

template <typename Real, bool NT, bool PROD = false>
__device__ __forceinline__ PairIO<Real, NT> make_pair_io(const LinePass& p, int32_t pair) {
  using C = cplx_t<Real>;
  PairIO<Real, NT> io;
  const int64_t ies = p.kind == kR2C ? sizeof(Real) : sizeof(C);
  const int64_t oes = p.kind == kC2R ? sizeof(Real) : sizeof(C);
  int32_t idx[kMaxDim];
  const int32_t la = 2 * pair;
  decode_line<Real>(p, la, idx);
  if constexpr (PROD) {
    io.ib[0] = product_source(p, idx, ies, io.is[0]);
  } else {
    io.ib[0] = element_ptr(p.in, idx, ies, p.axis, io.is[0]);
  }
  io.ob[0] = element_ptr(p.out, idx, oes, p.axis, io.os[0]);
  io.has_b = la + 1 < p.nlines;
  if (io.has_b) {
    decode_line<Real>(p, la + 1, idx);
    if constexpr (PROD) {
      io.ib[1] = io.ib[0];
      io.is[1] = io.is[0];
    } else {
      io.ib[1] = element_ptr(p.in, idx, ies, p.axis, io.is[1]);
    }
    io.ob[1] = element_ptr(p.out, idx, oes, p.axis, io.os[1]);
  } else {
    io.ib[1] = io.ib[0];
    io.ob[1] = io.ob[0];
    io.is[1] = io.is[0];
    io.os[1] = io.os[0];
  }
  io.kind = p.kind;
  io.n = p.n;
  io.scale = static_cast<Real>(p.scale);
  io.hw = p.herm_weight != 0.0 ? static_cast<Real>(p.herm_weight) : Real(1);
  return io;
}

// ---------------------------------------------------------------------------
// Compile-time-length Stockham kernel: each line is processed by T = N/R
// threads holding R values in registers; radix stages exchange through LDS.
// ---------------------------------------------------------------------------
template <typename Real> struct LdsPad;
template <> struct LdsPad<double> { static constexpr int shift = 3; };
template <> struct LdsPad<float> { static constexpr int shift = 4; };

template <typename Real, int N, int RM> struct FastCfg {
  static constexpr FPlan P = make_plan(N, RM);
  static constexpr int R = P.R;
  static constexpr int T = N / R;
  static constexpr int NST = P.nst;
  static constexpr int LDSN = N + (N >> LdsPad<Real>::shift) + 1;
  __device__ static __forceinline__ int lidx(int pos) { return pos + (pos >> LdsPad<Real>::shift); }
};

// Twiddles of a radix-r butterfly are the powers w^j (j = 1..r-1) of one
// table entry w = exp(-2 pi i k / (L r)): one FP64 load, the other powers by
// FP64 products (at most three multiplications deep), each rounded once to the
// kernel's precision.
template <int r>
__device__ __forceinline__ void twiddle_powers(double2 w, double2* p) {
  p[1] = w;
  if constexpr (r > 2) p[2] = cmul(w, w);
  if constexpr (r > 3) p[3] = cmul(p[2], w);
  if constexpr (r > 4) p[4] = cmul(p[2], p[2]);
  if constexpr (r > 5) p[5] = cmul(p[4], w);
  if constexpr (r > 6) p[6] = cmul(p[3], p[3]);
  if constexpr (r > 7) p[7] = cmul(p[4], p[3]);
}

template <typename C> __device__ __forceinline__ C to_prec(double2 w);
template <> __device__ __forceinline__ double2 to_prec<double2>(double2 w) { return w; }
template <> __device__ __forceinline__ float2 to_prec<float2>(double2 w) {
  return float2{static_cast<float>(w.x), static_cast<float>(w.y)};
}

// One twiddle-table entry, read as a native vector from global memory (a
// double2 is a class: copying it through a global pointer would pass a
// generic reference and compile to a flat load).
__device__ __forceinline__ double2 twiddle_load(const double2* p) {
  typedef double D2 __attribute__((ext_vector_type(2)));
  const D2 w = gload<false>(reinterpret_cast<const D2*>(p));
  return double2{w.x, w.y};
}

// The twiddle-table entries radix stage S of lane t multiplies by.  When a
// lane needs at most four (a few registers), they are loaded ahead: the
// stage loop issues them before the LDS exchange that precedes the stage, so
// their latency passes under the exchange's barriers instead of stalling
// every wave of the block right after them.
template <typename Real, int N, int RM, int S> struct StageTwiddles {
  using Cfg = FastCfg<Real, N, RM>;
  static constexpr int r = Cfg::P.rad[S];
  static constexpr int L = span_before(N, S);
  static constexpr int NB = Cfg::R / r;
  static constexpr int ts = N / (L * r);
  static constexpr bool kAhead = L > 1 && NB <= 4;
  // more entries per lane (the last radix-2 stage of 16 values per thread):
  // when NB divides L they are the consecutive entries k0 .. k0 + NB - 1,
  // k0 = t*NB mod L, and entry k0 + q is entry k0 times entry q, the same for
  // every lane (a scalar load): one entry per lane is loaded ahead
  static constexpr bool kBase = L > 1 && NB > 4 && L % NB == 0;
  double2 w[kAhead ? NB : 1];
  __device__ __forceinline__ void load(int t, const double2* tw) {
    if constexpr (kAhead) {
#pragma unroll
      for (int q = 0; q < NB; ++q) w[q] = twiddle_load(tw + ((t * NB + q) % L) * ts);
    } else if constexpr (kBase) {
      w[0] = twiddle_load(tw + ((t * NB) % L) * ts);
    }
  }
  __device__ __forceinline__ double2 get(int t, const double2* tw, int q) const {
    if constexpr (kAhead) return w[q];
    else if constexpr (kBase) return q == 0 ? w[0] : cmul(w[0], twiddle_load(tw + q * ts));
    else return twiddle_load(tw + ((t * NB + q) % L) * ts);
  }
};

template <typename Real, int N, int RM, int S>
__device__ __forceinline__ void stage_compute(cplx_t<Real>* v, int t, const double2* tw,
                                              const StageTwiddles<Real, N, RM, S>& pre) {
  using Cfg = FastCfg<Real, N, RM>;
  using C = cplx_t<Real>;
  constexpr int r = Cfg::P.rad[S];
  constexpr int L = span_before(N, S);
  constexpr int R = Cfg::R;
  constexpr int NB = R / r;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    if constexpr (L > 1) {
      double2 pw[8];
      twiddle_powers<r>(pre.get(t, tw, q), pw);
#pragma unroll
      for (int j = 1; j < r; ++j) v[q * r + j] = cmul(v[q * r + j], to_prec<C>(pw[j]));
    }
    bfly<r>(v + q * r);
  }
}

// Bytes of one line's LDS buffer: LDSN complex values, or LDSN reals when
// the exchanges travel in two halves (SPLIT).
template <typename Real, int N, int RM, bool SPLIT> constexpr size_t line_lds_bytes() {
  return static_cast<size_t>(FastCfg<Real, N, RM>::LDSN) * (SPLIT ? sizeof(Real) : sizeof(cplx_t<Real>));
}

// Line buffer of the two-lines-per-lane kernel (reals): LDSN plus a pad
// measured per length.  1024-point FP32 lines: one more float per line moves
// the 16 line buffers a wave touches onto different banks (lane pairs read
// and write 4 positions of each of 16 lines per access); the y pass of a
// 1024^3 R2C plan took 1.80 instead of 2.08 ms with it, while the same pad
// made 512-point lines 5-8 % slower.
template <typename Real, int N> constexpr int dual_lds_extra() {
  return (sizeof(Real) == 4 && N == 1024) ? 1 : 0;
}
template <typename Real, int N, int RM> constexpr size_t dual_line_bytes() {
  return static_cast<size_t>(FastCfg<Real, N, RM>::LDSN + dual_lds_extra<Real, N>()) * sizeof(Real);
}

// Register <-> LDS permutation of one line: every lane stores its values
// v[q*RW + j] at positions wpos(q, j), then loads v[q*RR + j] from positions
// rpos(q, j) (NW*RW = NR*RR = R).  SPLIT moves the real parts first and the
// imaginary parts second through a buffer of reals: after the first half a
// lane's old real parts are in LDS and can be overwritten, its old imaginary
// parts are still in registers for the second half.
template <typename Real, int N, int RM, bool SPLIT, int NW, int RW, int NR, int RR, typename WPos, typename RPos>
__device__ __forceinline__ void lds_permute(cplx_t<Real>* v, void* line, WPos wpos, RPos rpos) {
  using Cfg = FastCfg<Real, N, RM>;
  if constexpr (SPLIT) {
    Real* lr = static_cast<Real*>(line);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NW; ++q) {
#pragma unroll
        for (int j = 0; j < RW; ++j) lr[Cfg::lidx(wpos(q, j))] = h ? v[q * RW + j].y : v[q * RW + j].x;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NR; ++q) {
#pragma unroll
        for (int j = 0; j < RR; ++j) {
          const Real x = lr[Cfg::lidx(rpos(q, j))];
          if (h) {
            v[q * RR + j].y = x;
          } else {
            v[q * RR + j].x = x;
          }
        }
      }
    }
  } else {
    cplx_t<Real>* lc = static_cast<cplx_t<Real>*>(line);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NW; ++q) {
#pragma unroll
      for (int j = 0; j < RW; ++j) lc[Cfg::lidx(wpos(q, j))] = v[q * RW + j];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NR; ++q) {
#pragma unroll
      for (int j = 0; j < RR; ++j) v[q * RR + j] = lc[Cfg::lidx(rpos(q, j))];
    }
  }
}

// write the outputs of stage S, read the inputs of stage S+1
template <typename Real, int N, int RM, bool SPLIT, int S>
__device__ __forceinline__ void stage_exchange(cplx_t<Real>* v, int t, void* line) {
  using Cfg = FastCfg<Real, N, RM>;
  constexpr int r = Cfg::P.rad[S];
  constexpr int L = span_before(N, S);
  constexpr int R = Cfg::R;
  constexpr int NB = R / r;
  constexpr int r2 = Cfg::P.rad[S + 1];
  constexpr int NB2 = R / r2;
  lds_permute<Real, N, RM, SPLIT, NB, r, NB2, r2>(
      v, line,
      [t](int q, int j) {
        const int beta = t * NB + q;
        return (beta / L) * L * r + beta % L + j * L;
      },
      [t](int q, int j) { return t * NB2 + q + j * (N / r2); });
}

// after the last stage the outputs sit at positions beta + j*(N/rl)
// (beta = t*(R/rl) + q); regroup them to positions t + j*T
template <typename Real, int N, int RM, bool SPLIT>
__device__ __forceinline__ void final_regroup(cplx_t<Real>* v, int t, void* line) {
  using Cfg = FastCfg<Real, N, RM>;
  constexpr int R = Cfg::R, T = Cfg::T;
  constexpr int rl = Cfg::P.rad[Cfg::NST - 1];
  if constexpr (rl != R) {
    constexpr int NBl = R / rl;
    lds_permute<Real, N, RM, SPLIT, NBl, rl, 1, R>(
        v, line, [t](int q, int j) { return t * NBl + q + j * (N / rl); },
        [t](int, int j) { return t + j * T; });
  }
}

template <typename Real, int N, int RM, bool SPLIT, int S> struct StageLoop {
  __device__ static __forceinline__ void run(cplx_t<Real>* v, int t, const double2* tw, void* line,
                                             const StageTwiddles<Real, N, RM, S>& pre) {
    stage_compute<Real, N, RM, S>(v, t, tw, pre);
    if constexpr (S + 1 < FastCfg<Real, N, RM>::NST) {
      StageTwiddles<Real, N, RM, S + 1> next;
      next.load(t, tw);
      stage_exchange<Real, N, RM, SPLIT, S>(v, t, line);
      StageLoop<Real, N, RM, SPLIT, S + 1>::run(v, t, tw, line, next);
    }
  }
  __device__ static __forceinline__ void run(cplx_t<Real>* v, int t, const double2* tw, void* line) {
    StageTwiddles<Real, N, RM, S> pre;
    pre.load(t, tw);
    run(v, t, tw, line, pre);
  }
};

// ---------------------------------------------------------------------------
// Two lines per lane (FP32 lanes-across-lines passes whose consecutive lines
// are adjacent elements of one row): lane b of a block holds lines 2b and
// 2b+1, one 16-byte access moves element k of both, and the two register
// sets share twiddles and barriers.  B = 32 lines per block reads 256-byte
// row segments with 16 lanes across the lines; the LDS exchanges always move
// real and imaginary halves (2 lines x N + N/16 + 1 floats per lane pair).
// ---------------------------------------------------------------------------
template <typename Real, int N, int RM, int S>
__device__ __forceinline__ void stage_compute2(cplx_t<Real>* va, cplx_t<Real>* vb, int t, const double2* tw,
                                               const StageTwiddles<Real, N, RM, S>& pre) {
  using Cfg = FastCfg<Real, N, RM>;
  using C = cplx_t<Real>;
  constexpr int r = Cfg::P.rad[S];
  constexpr int L = span_before(N, S);
  constexpr int R = Cfg::R;
  constexpr int NB = R / r;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    if constexpr (L > 1) {
      double2 pw[8];
      twiddle_powers<r>(pre.get(t, tw, q), pw);
#pragma unroll
      for (int j = 1; j < r; ++j) {
        const C w = to_prec<C>(pw[j]);
        va[q * r + j] = cmul(va[q * r + j], w);
        vb[q * r + j] = cmul(vb[q * r + j], w);
      }
    }
    bfly<r>(va + q * r);
    bfly<r>(vb + q * r);
  }
}

// lds_permute for the two register sets of a lane (buffers la, lb of reals)
template <typename Real, int N, int RM, int NW, int RW, int NR, int RR, typename WPos, typename RPos>
__device__ __forceinline__ void lds_permute2(cplx_t<Real>* va, cplx_t<Real>* vb, Real* la, Real* lb, WPos wpos,
                                             RPos rpos) {
  using Cfg = FastCfg<Real, N, RM>;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NW; ++q) {
#pragma unroll
      for (int j = 0; j < RW; ++j) {
        const int i = Cfg::lidx(wpos(q, j));
        la[i] = h ? va[q * RW + j].y : va[q * RW + j].x;
        lb[i] = h ? vb[q * RW + j].y : vb[q * RW + j].x;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NR; ++q) {
#pragma unroll
      for (int j = 0; j < RR; ++j) {
        const int i = Cfg::lidx(rpos(q, j));
        const Real xa = la[i], xb = lb[i];
        if (h) {
          va[q * RR + j].y = xa;
          vb[q * RR + j].y = xb;
        } else {
          va[q * RR + j].x = xa;
          vb[q * RR + j].x = xb;
        }
      }
    }
  }
}

template <typename Real, int N, int RM, int S> struct StageLoop2 {
  __device__ static __forceinline__ void run(cplx_t<Real>* va, cplx_t<Real>* vb, int t, const double2* tw, Real* la,
                                             Real* lb) {
    StageTwiddles<Real, N, RM, S> pre;
    pre.load(t, tw);
    run(va, vb, t, tw, la, lb, pre);
  }
  __device__ static __forceinline__ void run(cplx_t<Real>* va, cplx_t<Real>* vb, int t, const double2* tw, Real* la,
                                             Real* lb, const StageTwiddles<Real, N, RM, S>& pre) {
    using Cfg = FastCfg<Real, N, RM>;
    stage_compute2<Real, N, RM, S>(va, vb, t, tw, pre);
    if constexpr (S + 1 < Cfg::NST) {
      StageTwiddles<Real, N, RM, S + 1> next;
      next.load(t, tw);
      constexpr int r = Cfg::P.rad[S];
      constexpr int L = span_before(N, S);
      constexpr int R = Cfg::R;
      constexpr int NB = R / r;
      constexpr int r2 = Cfg::P.rad[S + 1];
      constexpr int NB2 = R / r2;
      lds_permute2<Real, N, RM, NB, r, NB2, r2>(
          va, vb, la, lb,
          [t](int q, int j) {
            const int beta = t * NB + q;
            return (beta / L) * L * r + beta % L + j * L;
          },
          [t](int q, int j) { return t * NB2 + q + j * (N / r2); });
      StageLoop2<Real, N, RM, S + 1>::run(va, vb, t, tw, la, lb, next);
    }
  }
};

// SPLIT_IN: the input side has odd rows under a padded enumeration; its
// line pairs are loaded as two 8-byte elements, and the second element of a
// pair whose second line is padding (the next row's first element there, or
// past the end of the rank's array) is not loaded.
template <typename Real, int N, int RM, bool NT, int SPLIT_IN = 0>
__device__ __forceinline__ void fast_line2_body(const LinePass& p, int32_t blk, int B, unsigned char* lds) {
  using C = cplx_t<Real>;
  using Cfg = FastCfg<Real, N, RM>;
  typedef Real V4 __attribute__((ext_vector_type(4)));
  constexpr int R = Cfg::R, T = Cfg::T;
  constexpr size_t LB = dual_line_bytes<Real, N, RM>();
  const int P2 = B >> 1;
  const int tid = threadIdx.x;
  // lanes beyond P2*T (grouped launches of 256 lanes) run the stages on zeros
  // in two spare LDS slots (2*P2, 2*P2 + 1) and load and store nothing
  const bool active = tid < P2 * T;
  const int b = active ? tid % P2 : P2, t = active ? tid / P2 : 0;
  const int32_t la = blk * B + 2 * b;  // lines la, la + 1 (nlines is even)
  const bool valid = active && la < p.nlines && !padding_line<Real>(p, la);
  const double2* tw = reinterpret_cast<const double2*>(p.tw);
  Real* ba = reinterpret_cast<Real*>(lds + (2 * b) * LB);
  Real* bb = reinterpret_cast<Real*>(lds + (2 * b + 1) * LB);
  LineIO<Real, NT> io;
  if (valid) io = make_line_io<Real, NT>(p, la);

  bool second = false;
  if constexpr (SPLIT_IN == 1) second = valid && !padding_line<Real>(p, la + 1);
  // SPLIT_IN == 2: line la + 1 has addresses of its own on both sides
  [[maybe_unused]] LineIO<Real, NT> io2 = io;
  if constexpr (SPLIT_IN == 2) {
    second = valid && la + 1 < p.nlines && !padding_line<Real>(p, la + 1);
    if (second) io2 = make_line_io<Real, NT>(p, la + 1);
  }
  C va[R], vb[R];
  constexpr int r0 = Cfg::P.rad[0];
  constexpr int NB0 = R / r0;
#pragma unroll
  for (int q = 0; q < NB0; ++q) {
    const C* src = reinterpret_cast<const C*>(io.ib) + static_cast<int64_t>(t * NB0 + q) * io.is;
    const int64_t step = static_cast<int64_t>(N / r0) * io.is;
    [[maybe_unused]] const C* src2 = nullptr;
    [[maybe_unused]] int64_t step2 = 0;
    if constexpr (SPLIT_IN == 2) {
      src2 = reinterpret_cast<const C*>(io2.ib) + static_cast<int64_t>(t * NB0 + q) * io2.is;
      step2 = static_cast<int64_t>(N / r0) * io2.is;
    }
#pragma unroll
    for (int j = 0; j < r0; ++j) {
      C a, c;
      if constexpr (SPLIT_IN == 2) {
        a = valid ? gload_c<NT, Real>(src) : C{Real(0), Real(0)};
        c = second ? gload_c<NT, Real>(src2) : C{Real(0), Real(0)};
        src2 += step2;
      } else if constexpr (SPLIT_IN == 1) {
        a = valid ? gload_c<NT, Real>(src) : C{Real(0), Real(0)};
        c = second ? gload_c<NT, Real>(src + 1) : C{Real(0), Real(0)};
      } else {
        const V4 x = valid ? gload<NT>(reinterpret_cast<const V4*>(src)) : V4{Real(0), Real(0), Real(0), Real(0)};
        a = C{x.x, x.y};
        c = C{x.z, x.w};
      }
      va[q * r0 + j] = p.inverse ? cconj(a) : a;
      vb[q * r0 + j] = p.inverse ? cconj(c) : c;
      src += step;
    }
  }
  StageLoop2<Real, N, RM, 0>::run(va, vb, t, tw, ba, bb);
  constexpr int rl = Cfg::P.rad[Cfg::NST - 1];
  if constexpr (rl != R) {
    constexpr int NBl = R / rl;
    lds_permute2<Real, N, RM, NBl, rl, 1, R>(
        va, vb, ba, bb, [t](int q, int j) { return t * NBl + q + j * (N / rl); },
        [t](int, int j) { return t + j * T; });
  }
  // SPLIT_IN == 2: the pair is stored as one 16-byte element pair where the
  // second line's output follows the first's on a 16-byte boundary at every
  // position (even element stride), else as two elements
  bool pair_out = true;
  if constexpr (SPLIT_IN == 2)
    pair_out = second && io2.ob == io.ob + sizeof(C) && io2.os == io.os && (io.os & 1) == 0 &&
               (reinterpret_cast<uintptr_t>(io.ob) & 15) == 0;
  if (valid && pair_out) {
    C* dst = reinterpret_cast<C*>(io.ob) + static_cast<int64_t>(t) * io.os;
    const int64_t step = static_cast<int64_t>(T) * io.os;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      C a = p.inverse ? cconj(va[j]) : va[j];
      C c = p.inverse ? cconj(vb[j]) : vb[j];
      V4 x;
      x.x = a.x * io.scale;
      x.y = a.y * io.scale;
      x.z = c.x * io.scale;
      x.w = c.y * io.scale;
      gstore<NT>(reinterpret_cast<V4*>(dst), x);
      dst += step;
    }
  } else if constexpr (SPLIT_IN == 2) {
    if (valid) {
      C* dst = reinterpret_cast<C*>(io.ob) + static_cast<int64_t>(t) * io.os;
      const int64_t step = static_cast<int64_t>(T) * io.os;
      [[maybe_unused]] C* dst2 = nullptr;
      [[maybe_unused]] int64_t stepb = 0;
      if (second) {
        dst2 = reinterpret_cast<C*>(io2.ob) + static_cast<int64_t>(t) * io2.os;
        stepb = static_cast<int64_t>(T) * io2.os;
      }
#pragma unroll
      for (int j = 0; j < R; ++j) {
        C a = p.inverse ? cconj(va[j]) : va[j];
        gstore_c<NT, Real>(dst, C{a.x * io.scale, a.y * io.scale});
        dst += step;
        if (second) {
          C c = p.inverse ? cconj(vb[j]) : vb[j];
          gstore_c<NT, Real>(dst2, C{c.x * io.scale, c.y * io.scale});
          dst2 += stepb;
        }
      }
    }
  }
}

// One block's share of a line pass: lines [blk*B, blk*B + B) of p, T = N/R
// lanes per line.  Lanes beyond B*T (a block shape shared by several line
// lengths) run the stages on zeros in a spare LDS slot (index B) so every
// lane reaches every barrier; they load and store nothing.  STAGE (complex64
// grouped passes) moves contiguous lines 16 bytes per lane through LDS.
template <typename Real, int N, int RM, bool NT, bool SPLIT, bool STAGE = false>
__device__ __forceinline__ void fast_line_body(const LinePass& p, int32_t blk, int B, unsigned char* lds) {
  using C = cplx_t<Real>;
  using Cfg = FastCfg<Real, N, RM>;
  constexpr int R = Cfg::R, T = Cfg::T;
  const int tid = threadIdx.x;
  int b, t;
  if (p.lines_fast) {
    b = tid % B;
    t = tid / B;
  } else {
    t = tid % T;
    b = tid / T;
  }
  const bool active = tid < B * T;
  if (!active) {
    b = B;
    t = 0;
  }
  const int32_t line = blk * B + b;
  const double2* tw = reinterpret_cast<const double2*>(p.tw);
  void* buf = lds + b * line_lds_bytes<Real, N, RM, SPLIT>();

  bool valid;
  LineIO<Real, NT> io;
  if (T >= 64 && !p.lines_fast) {
    // lanes along the line with T >= 64: each wave holds one line (or only
    // spare lanes), so the line index is wave-uniform and the descriptor
    // arithmetic of make_line_io runs once per wave on the scalar unit
    const int32_t wl = __builtin_amdgcn_readfirstlane(line);
    valid = active && wl < p.nlines && !padding_line<Real>(p, wl);
    if (valid) io = make_line_io<Real, NT>(p, wl);
  } else {
    valid = active && line < p.nlines && !padding_line<Real>(p, line);
    if (valid) io = make_line_io<Real, NT>(p, line);
  }

  C v[R];
  constexpr int r0 = Cfg::P.rad[0];
  constexpr int NB0 = R / r0;
  typedef Real V4 __attribute__((ext_vector_type(4)));
  static_assert(!STAGE || (std::is_same<Real, float>::value && !SPLIT && R % 2 == 0),
                "staged lines are complex64 with complex LDS exchanges");
  bool vin = false;
  if constexpr (STAGE) {
    // contiguous C2C lines that start on a 16-byte boundary are read as
    // element pairs (lane t takes pairs t, t + T, ...: T lanes cover 16*T
    // adjacent bytes per access) and regrouped through the line's LDS buffer
    C* lc = static_cast<C*>(buf);
    vin = valid && p.kind == kC2C && io.is == 1 && (reinterpret_cast<uintptr_t>(io.ib) & 15) == 0;
    if (vin) {
      const V4* src = reinterpret_cast<const V4*>(io.ib) + t;
#pragma unroll
      for (int k = 0; k < R / 2; ++k) {
        const V4 x = gload<NT>(src + k * T);
        const int e = 2 * (t + k * T);
        lc[Cfg::lidx(e)] = C{x.x, x.y};
        lc[Cfg::lidx(e + 1)] = C{x.z, x.w};
      }
    }
    if (p.kind == kC2C) __syncthreads();  // p is the same member for the whole block
    if (vin) {
#pragma unroll
      for (int q = 0; q < NB0; ++q) {
#pragma unroll
        for (int j = 0; j < r0; ++j) {
          const C x = lc[Cfg::lidx(t * NB0 + q + j * (N / r0))];
          v[q * r0 + j] = p.inverse ? cconj(x) : x;
        }
      }
    }
  }
  if (vin) {
    // loaded above, through LDS
  } else if (p.kind == kC2C) {
    // C2C lines are walked with pointer increments (no 64-bit product per
    // element).  One branch around all the loads: a select per load puts
    // each load under its own exec mask, and the zeroing moves between them
    // made the compiler wait for the loads in flight halfway through.
    if (valid) {
#pragma unroll
      for (int q = 0; q < NB0; ++q) {
        const C* src = reinterpret_cast<const C*>(io.ib) + static_cast<int64_t>(t * NB0 + q) * io.is;
        const int64_t step = static_cast<int64_t>(N / r0) * io.is;
#pragma unroll
        for (int j = 0; j < r0; ++j) {
          v[q * r0 + j] = gload_c<NT, Real>(src);
          src += step;
        }
      }
      if (p.inverse) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = cconj(v[j]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = C{Real(0), Real(0)};
    }
  } else if (valid) {
    // R2C / C2R lines (one branch around the loads, as for C2C, and one per kind)
    if (p.kind == kR2C) {
#pragma unroll
      for (int q = 0; q < NB0; ++q) {
#pragma unroll
        for (int j = 0; j < r0; ++j) v[q * r0 + j] = io.template load_kind<kR2C>(t * NB0 + q + j * (N / r0));
      }
    } else {
#pragma unroll
      for (int q = 0; q < NB0; ++q) {
#pragma unroll
        for (int j = 0; j < r0; ++j) v[q * r0 + j] = io.template load_kind<kC2R>(t * NB0 + q + j * (N / r0));
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = C{Real(0), Real(0)};
  }
  StageLoop<Real, N, RM, SPLIT, 0>::run(v, t, tw, buf);
  bool vout = false;
  if constexpr (STAGE) {
    // the last stage's outputs go to LDS in natural order (this replaces
    // final_regroup); contiguous aligned C2C outputs are stored as element
    // pairs, any other line reads back positions t + j*T
    constexpr int rl = Cfg::P.rad[Cfg::NST - 1];
    constexpr int NBl = R / rl;
    C* lc = static_cast<C*>(buf);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NBl; ++q) {
#pragma unroll
      for (int j = 0; j < rl; ++j) lc[Cfg::lidx(t * NBl + q + j * (N / rl))] = v[q * rl + j];
    }
    __syncthreads();
    vout = valid && p.kind == kC2C && io.os == 1 && (reinterpret_cast<uintptr_t>(io.ob) & 15) == 0;
    if (vout) {
      V4* dst = reinterpret_cast<V4*>(io.ob) + t;
#pragma unroll
      for (int k = 0; k < R / 2; ++k) {
        const int e = 2 * (t + k * T);
        C a = lc[Cfg::lidx(e)], c = lc[Cfg::lidx(e + 1)];
        if (p.inverse) {
          a = cconj(a);
          c = cconj(c);
        }
        V4 x;
        x.x = a.x * io.scale;
        x.y = a.y * io.scale;
        x.z = c.x * io.scale;
        x.w = c.y * io.scale;
        gstore<NT>(dst + k * T, x);
      }
    } else {
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = lc[Cfg::lidx(t + j * T)];
    }
  } else {
    final_regroup<Real, N, RM, SPLIT>(v, t, buf);
  }

  if (valid && !vout) {
    if (p.kind == kC2C) {
      C* dst = reinterpret_cast<C*>(io.ob) + static_cast<int64_t>(t) * io.os;
      const int64_t step = static_cast<int64_t>(T) * io.os;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        C x = p.inverse ? cconj(v[j]) : v[j];
        x.x *= io.scale;
        x.y *= io.scale;
        gstore_c<NT, Real>(dst, x);
        dst += step;
      }
    } else {
#pragma unroll
      for (int j = 0; j < R; ++j) io.store(t + j * T, v[j]);
    }
  }
}

// Paired R2C / C2R: slot b of the block transforms lines 2q and 2q+1
// (q = blk*B + b) as one complex line; same stage code as fast_line_body.
// PROD: a products pass (LinePass::products): each loaded real value is the
// product of the two factor lines' values at that position.
template <typename Real, int N, int RM, bool NT, bool SPLIT, bool PROD = false>
__device__ __forceinline__ void fast_pair_body(const LinePass& p, int32_t blk, int B, unsigned char* lds) {
  using C = cplx_t<Real>;
  using Cfg = FastCfg<Real, N, RM>;
  constexpr int R = Cfg::R, T = Cfg::T;
  const int tid = threadIdx.x;
  int b, t;
  if (p.lines_fast) {
    b = tid % B;
    t = tid / B;
  } else {
    t = tid % T;
    b = tid / T;
  }
  const bool active = tid < B * T;
  if (!active) {
    b = B;
    t = 0;
  }
  const int32_t pair = blk * B + b;
  const bool valid = active && 2 * pair < p.nlines;  // paired passes never enumerate padding
  const double2* tw = reinterpret_cast<const double2*>(p.tw);
  void* buf = lds + b * line_lds_bytes<Real, N, RM, SPLIT>();

  PairIO<Real, NT> io{};
  if (T >= 64 && !p.lines_fast) {
    // lanes along the pair with T >= 64: each wave holds one pair (or only
    // spare lanes), so the pair index is wave-uniform and the descriptor
    // arithmetic of make_pair_io runs once per wave on the scalar unit
    const int32_t wp = __builtin_amdgcn_readfirstlane(pair);
    if (valid) io = make_pair_io<Real, NT, PROD>(p, wp);
  } else {
    if (valid) io = make_pair_io<Real, NT, PROD>(p, pair);
  }

  // Products pass: the block holds whole positions (B is a multiple of 3 and
  // the field axis is enumerated fastest, so slots 3m..3m+2 are the pairs
  // (0,1), (2,3), (4,5) of one position).  Slot 3m+q stages velocity line q of
  // the position into the LDS of the position's three slots; each slot then
  // forms its two product lines from LDS.  Global loads: 3 lines per
  // position instead of 2 per product line.
  [[maybe_unused]] const Real* stage = nullptr;
  [[maybe_unused]] int pq = 0;
  if constexpr (PROD) {
    pq = b % 3;
    Real* st = reinterpret_cast<Real*>(lds + (b - pq) * line_lds_bytes<Real, N, RM, SPLIT>());
    if (valid) {
      const Real* src = reinterpret_cast<const Real*>(io.ib[0]) + static_cast<int64_t>(t) * io.is[0];
      const int64_t step = static_cast<int64_t>(T) * io.is[0];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        st[pq * N + t + j * T] = gload<NT>(src);
        src += step;
      }
    }
    __syncthreads();
    stage = st;
  }

  // line elements are walked with pointer increments (positions beta + j*S
  // on load, t + j*T on store): no 64-bit product per element
  C v[R];
  C mirror[R];  // C2R: input at position n - pos, formed by the lane that loads X[pos]
  {
    constexpr int r0 = Cfg::P.rad[0];
    constexpr int NB0 = R / r0;
    constexpr int S = N / r0;
    const bool vb = valid && io.has_b;
    const int64_t sa = static_cast<int64_t>(S) * io.is[0], sb = static_cast<int64_t>(S) * io.is[1];
#pragma unroll
    for (int q = 0; q < NB0; ++q) {
      const int beta = t * NB0 + q;
      if (p.kind == kR2C) {
        const Real* pa = reinterpret_cast<const Real*>(io.ib[0]) + static_cast<int64_t>(beta) * io.is[0];
        const Real* pb = reinterpret_cast<const Real*>(io.ib[1]) + static_cast<int64_t>(beta) * io.is[1];
#pragma unroll
        for (int j = 0; j < r0; ++j) {
          Real xa, xb;
          if constexpr (PROD) {
            const int pos = beta + j * S;
            const int fa = 2 * pq, fb = 2 * pq + 1;
            xa = valid ? stage[prod_first(fa) * N + pos] * stage[prod_second(fa) * N + pos] : Real(0);
            xb = vb ? stage[prod_first(fb) * N + pos] * stage[prod_second(fb) * N + pos] : Real(0);
          } else {
            xa = valid ? gload<NT>(pa) : Real(0);
            xb = vb ? gload<NT>(pb) : Real(0);
          }
          v[q * r0 + j] = C{xa, xb};
          pa += sa;
          pb += sb;
        }
      } else {
        // C2R: each half-spectrum entry X[m], m <= n/2, is loaded once, by the
        // lane holding position m; that lane also forms the input at the
        // mirrored position n - m (from conj X[m]) and hands it over through
        // LDS below
        // beta < S, so positions beta + j*S with j < r0/2 are below n/2, the
        // one with j = r0/2 reaches n/2 only for beta = 0 and the others lie
        // above: the loads of the first r0/2 positions are unconditional and
        // sit under one branch per line (a branch per load would wait for
        // each pair of loads before the next)
        constexpr int JH = r0 / 2;
        const C* la = reinterpret_cast<const C*>(io.ib[0]) + static_cast<int64_t>(beta) * io.is[0];
        const C* lb = reinterpret_cast<const C*>(io.ib[1]) + static_cast<int64_t>(beta) * io.is[1];
        const bool mid = 2 * (beta + JH * S) <= N;
        C A[JH + 1], Bv[JH + 1];
#pragma unroll
        for (int j = 0; j <= JH; ++j) {
          A[j] = C{Real(0), Real(0)};
          Bv[j] = C{Real(0), Real(0)};
        }
        if (valid) {
#pragma unroll
          for (int j = 0; j < JH; ++j) A[j] = gload_c<NT, Real>(la + j * sa);
          if (mid) A[JH] = gload_c<NT, Real>(la + JH * sa);
        }
        if (vb) {
#pragma unroll
          for (int j = 0; j < JH; ++j) Bv[j] = gload_c<NT, Real>(lb + j * sb);
          if (mid) Bv[JH] = gload_c<NT, Real>(lb + JH * sb);
        }
#pragma unroll
        for (int j = 0; j <= JH; ++j) {
          const int pos = beta + j * S;
          if (j < JH || mid) {
            v[q * r0 + j] = io.c2r_value(pos, true, A[j], Bv[j]);
            mirror[q * r0 + j] = io.c2r_value(N - pos, false, A[j], Bv[j]);
          }
        }
      }
    }
    if (p.kind == kC2R) {
      // positions above n/2 receive the value formed by the lane holding n - pos
      if constexpr (SPLIT) {
        Real* lr = static_cast<Real*>(buf);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (h) __syncthreads();
#pragma unroll
          for (int q = 0; q < NB0; ++q) {
#pragma unroll
            for (int j = 0; j < r0; ++j) {
              const int pos = t * NB0 + q + j * S;
              if (pos > 0 && 2 * pos < N) lr[Cfg::lidx(N - pos)] = h ? mirror[q * r0 + j].y : mirror[q * r0 + j].x;
            }
          }
          __syncthreads();
#pragma unroll
          for (int q = 0; q < NB0; ++q) {
#pragma unroll
            for (int j = 0; j < r0; ++j) {
              const int pos = t * NB0 + q + j * S;
              if (2 * pos > N) {
                const Real x = lr[Cfg::lidx(pos)];
                if (h) {
                  v[q * r0 + j].y = x;
                } else {
                  v[q * r0 + j].x = x;
                }
              }
            }
          }
        }
      } else {
        C* lc = static_cast<C*>(buf);
#pragma unroll
        for (int q = 0; q < NB0; ++q) {
#pragma unroll
          for (int j = 0; j < r0; ++j) {
            const int pos = t * NB0 + q + j * S;
            if (pos > 0 && 2 * pos < N) lc[Cfg::lidx(N - pos)] = mirror[q * r0 + j];
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NB0; ++q) {
#pragma unroll
          for (int j = 0; j < r0; ++j) {
            const int pos = t * NB0 + q + j * S;
            if (2 * pos > N) v[q * r0 + j] = lc[Cfg::lidx(pos)];
          }
        }
      }
    }
  }
  if constexpr (PROD) __syncthreads();  // the staged lines share LDS with the exchanges
  StageLoop<Real, N, RM, SPLIT, 0>::run(v, t, tw, buf);
  final_regroup<Real, N, RM, SPLIT>(v, t, buf);
  const int64_t ta = static_cast<int64_t>(T) * io.os[0], tb = static_cast<int64_t>(T) * io.os[1];
  if (p.kind == kC2R) {
    // w = FFT(conj Z) at each position, so a + i b = conj(w)
    if (valid) {
      Real* pa = reinterpret_cast<Real*>(io.ob[0]) + static_cast<int64_t>(t) * io.os[0];
      Real* pb = reinterpret_cast<Real*>(io.ob[1]) + static_cast<int64_t>(t) * io.os[1];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        gstore<NT>(pa, v[j].x * io.scale);
        if (io.has_b) gstore<NT>(pb, -v[j].y * io.scale);
        pa += ta;
        pb += tb;
      }
    }
    return;
  }
  // R2C: Z_{n-pos} lives in another lane; exchange the whole line through LDS
  C w[R];  // w[j] = Z_{(n - pos) mod n} for pos = t + j*T
  if constexpr (SPLIT) {
    Real* lr = static_cast<Real*>(buf);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < R; ++j) lr[Cfg::lidx(t + j * T)] = h ? v[j].y : v[j].x;
      __syncthreads();
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int pos = t + j * T;
        const Real x = lr[Cfg::lidx(pos == 0 ? 0 : N - pos)];
        if (h) {
          w[j].y = x;
        } else {
          w[j].x = x;
        }
      }
    }
  } else {
    C* lc = static_cast<C*>(buf);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) lc[Cfg::lidx(t + j * T)] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int pos = t + j * T;
      w[j] = lc[Cfg::lidx(pos == 0 ? 0 : N - pos)];
    }
  }
  if (valid) {
    C* pa = reinterpret_cast<C*>(io.ob[0]) + static_cast<int64_t>(t) * io.os[0];
    C* pb = reinterpret_cast<C*>(io.ob[1]) + static_cast<int64_t>(t) * io.os[1];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int pos = t + j * T;
      if (2 * pos <= N) {
        C A, Bv;
        io.r2c_values(pos, v[j], w[j], A, Bv);
        gstore_c<NT, Real>(pa, A);
        if (io.has_b) gstore_c<NT, Real>(pb, Bv);
      }
      pa += ta;
      pb += tb;
    }
  }
}

// Workgroups go to the 8 XCDs round-robin; renumber them so XCD x owns the
// x-th contiguous range of blocks (a bijection for any grid size).
__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t nb) {
  constexpr int32_t kXcds = 8;
  const int32_t q = nb / kXcds, r = nb % kXcds;
  const int32_t x = b % kXcds, i = b / kXcds;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Block of the pass for hardware block b: optional XCD-contiguous numbering,
// then an optional column-major walk of the (nb / C) x C block grid, where C
// blocks cover one row of the fastest batch axis (p.swizzle = C; the launcher
// sets it only when it divides the grid).
__device__ __forceinline__ int32_t pass_block(const LinePass& p, int32_t b, int32_t nb) {
  if (p.xcd_remap) b = xcd_block(b, nb);
  if (p.swizzle > 1) {
    const int32_t rows = nb / p.swizzle;
    b = (b % rows) * p.swizzle + b / rows;
  }
  return b;
}

template <typename Real, int N, int RM, bool NT, bool SPLIT, bool PROD = false>
__global__ void __launch_bounds__(1024) fft_pair_fast(const LinePass p) {
  extern __shared__ unsigned char smem_raw[];
  const int32_t blk = pass_block(p, blockIdx.x, gridDim.x);
  fast_pair_body<Real, N, RM, NT, SPLIT, PROD>(p, blk, p.lines_per_block, smem_raw);
}

template <typename Real, int N, int RM, bool NT, bool SPLIT>
__global__ void __launch_bounds__(1024) fft_line_fast(const LinePass p) {
  extern __shared__ unsigned char smem_raw[];
  const int32_t blk = pass_block(p, blockIdx.x, gridDim.x);
  fast_line_body<Real, N, RM, NT, SPLIT>(p, blk, p.lines_per_block, smem_raw);
}

template <typename Real, int N, int RM, bool NT, int SPLIT_IN>
__global__ void __launch_bounds__(1024) fft_line2_fast(const LinePass p) {
  extern __shared__ unsigned char smem_raw[];
  const int32_t blk = pass_block(p, blockIdx.x, gridDim.x);
  fast_line2_body<Real, N, RM, NT, SPLIT_IN>(p, blk, p.lines_per_block, smem_raw);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
namespace {

// Raises kernel fn's dynamic-LDS limit to at least `bytes` (the cached
// limit belongs to that one kernel; updates only grow it and are serialised,
// so concurrent launches from several host threads never lower it).  Never
// called inside a stream capture that follows an eager launch of the same
// shape: the limit is then already raised.
inline hipError_t raise_lds_limit(std::atomic<size_t>& limit, const void* fn, size_t bytes) {
  if (bytes <= limit.load(std::memory_order_acquire)) return hipSuccess;
  static std::mutex m;
  std::lock_guard<std::mutex> lock(m);
  if (bytes <= limit.load(std::memory_order_relaxed)) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
  if (e != hipSuccess) return e;
  limit.store(bytes, std::memory_order_release);
  return hipSuccess;
}

// One compile-time variant of the fast kernels: paired or single lines,
// nontemporal or cached accesses, split or complex LDS exchanges.
template <typename Real, int N, int RM, bool NT, bool SPLIT> struct FastVariant {
  static constexpr bool kPairs = !(std::is_same<Real, double>::value && (N % 3 == 0 || N % 5 == 0));
  static const void* fn(bool pair) {
    if constexpr (kPairs) {
      if (pair) return reinterpret_cast<const void*>(&fft_pair_fast<Real, N, 8, NT, SPLIT>);
    }
    return reinterpret_cast<const void*>(&fft_line_fast<Real, N, RM, NT, SPLIT>);
  }
  static void launch(bool pair, unsigned grid, unsigned threads, size_t smem, hipStream_t s, const LinePass& p) {
    if constexpr (kPairs) {
      if (pair) {
        hipLaunchKernelGGL((fft_pair_fast<Real, N, 8, NT, SPLIT>), dim3(grid), dim3(threads), smem, s, p);
        return;
      }
    }
    hipLaunchKernelGGL((fft_line_fast<Real, N, RM, NT, SPLIT>), dim3(grid), dim3(threads), smem, s, p);
  }
};

// split exchanges exist for FP64, and for FP32 lines of 1024 points and more
// (shorter FP32 lines are small in LDS already and ran slower split)
template <typename Real, int N> constexpr bool split_available() {
  return std::is_same<Real, double>::value || N >= 1024;
}

template <typename Real, int N, int RM, bool NT>
void launch_fast_variant(bool split, bool pair, unsigned grid, unsigned threads, size_t smem, hipStream_t s,
                         const LinePass& p) {
  if constexpr (split_available<Real, N>()) {
    if (split) {
      FastVariant<Real, N, RM, NT, true>::launch(pair, grid, threads, smem, s, p);
      return;
    }
  }
  FastVariant<Real, N, RM, NT, false>::launch(pair, grid, threads, smem, s, p);
}

template <typename Real, int N, int RM>
const void* fast_variant_fn(bool split, bool pair, bool nt) {
  if constexpr (split_available<Real, N>()) {
    if (split) return nt ? FastVariant<Real, N, RM, true, true>::fn(pair) : FastVariant<Real, N, RM, false, true>::fn(pair);
  }
  return nt ? FastVariant<Real, N, RM, true, false>::fn(pair) : FastVariant<Real, N, RM, false, false>::fn(pair);
}

// Two lines per lane applies to FP32 C2C lanes-across-lines passes of
// power-of-two lengths >= 128 whose consecutive lines are adjacent elements
// of one row on both sides (the fastest batch axis is the innermost storage
// axis, not split over ranks, with even row lengths (pitches) and 16-byte
// aligned rank buffers, so every line pair is one aligned 16-byte element
// pair), with an even number of lines per block, of lines and of the first
// line.
template <typename Real, int N> constexpr bool dual_kernel_exists() {
  return std::is_same<Real, float>::value && N >= 128 && (N & (N - 1)) == 0;
}

// Returns -1 where two lines per lane do not apply, else the dual_split
// mask of the pass.  An input side whose rows are odd is accepted when the
// lines are enumerated over a padded extent of even pitch along the row axis
// (pairs then never straddle two rows): its pairs are loaded as two 8-byte
// elements and the padding half of a row's last pair is skipped.  (Storing
// such a side as two 8-byte elements measured slower than the one-line
// kernel on the 1024^3 x pass, so odd output rows keep one line per lane.)
inline int dual_lines_mode(const LinePass& p, int B) {
  if (!p.dual_lines || p.kind != kC2C || !p.lines_fast || p.nbatch < 1 || B < 2 || B % 2) return -1;
  const int j = p.nbatch - 1, fa = p.batch_axis[j];
  // an odd unpadded row length of a padded array is fine: the pair at the
  // end of a row then covers one padding element, which lies inside the row
  // pitch and is never read by any pass
  if (p.batch_ext[j] % 2 || p.line_begin % 2 || p.nlines % 2) return -1;
  const bool padded_rows = p.pad_real > 0 && p.pad_axis == fa;
  int mask = 0, bit = 1;
  for (const DDesc* d : {&p.in, &p.out}) {
    const int inner = d->layout_left ? d->map[0] : d->map[d->ndim - 1];
    if (inner != fa || d->ntopo[fa] != 1) return -1;
    const int32_t row = (!d->layout_left && d->pitch > 0) ? d->pitch : d->ext[fa];
    int ranks = 1;
    for (int a = 0; a < d->ndim; ++a) ranks *= d->ntopo[a];
    bool aligned = true;
    for (int r = 0; r < ranks && r < kMaxRanks; ++r)
      if (reinterpret_cast<uintptr_t>(d->ptr[r]) % 16 != 0) aligned = false;
    if (row % 2) {
      if (!padded_rows || bit != 1) return -1;
      mask |= bit;
    } else if (!aligned) {
      return -1;
    }
    bit <<= 1;
  }
  return mask;
}

inline bool dual_lines_apply_impl(const LinePass& p, int B) { return dual_lines_mode(p, B) == 0; }

// Where dual_lines_mode refuses odd output rows: an unpadded output whose
// rows run along the fastest batch axis, both sides with that axis innermost
// and undistributed.  Enumerated over the unpadded rows, lines 2b and 2b+1
// are consecutive output elements, also across a row end; the kernel then
// takes each line's addresses on their own (two-lines mode 2) and pairs the
// stores where they are adjacent and 16-byte aligned.  Returns false or
// rewrites q into that enumeration.
// A lanes-across-lines C2C pass whose output rows (the fastest batch axis)
// are unpadded and odd, read from a scratch with padded rows: its lines are
// enumerated over the real elements only, so a block's consecutive lines are
// consecutive output elements also across a row end (whole, aligned output
// segments; each line takes its input address on its own).  False if the
// input rows are not padded (nothing to change).
inline bool lines_linear_unpadded(const LinePass& p, LinePass& q) {
  if (p.kind != kC2C || !p.lines_fast || p.nbatch < 1 || p.line_begin != 0 || p.pad_real <= 0) return false;
  const int j = p.nbatch - 1, fa = p.batch_axis[j];
  for (const DDesc* d : {&p.in, &p.out}) {
    if (d->layout_left) return false;
    if (d->map[d->ndim - 1] != fa || d->ntopo[fa] != 1) return false;
  }
  if (p.out.pitch > 0 || p.out.ext[fa] % 2 == 0) return false;
  if (p.pad_axis != fa || p.batch_ext[j] <= 0) return false;
  q = p;
  q.nlines = p.nlines / p.batch_ext[j] * p.pad_real;
  q.batch_ext[j] = p.pad_real;
  q.batch_div[j] = FastDiv(static_cast<uint32_t>(p.pad_real));
  q.pad_real = 0;
  q.pad_axis = 0;
  return true;
}

inline bool dual_lines_linear(const LinePass& p, int B, LinePass& q) {
  if (!p.dual_lines || p.kind != kC2C || !p.lines_fast || p.nbatch < 1 || B < 2 || B % 2) return false;
  if (p.line_begin != 0) return false;
  const int j = p.nbatch - 1, fa = p.batch_axis[j];
  for (const DDesc* d : {&p.in, &p.out}) {
    if (d->layout_left) return false;
    if (d->map[d->ndim - 1] != fa || d->ntopo[fa] != 1) return false;
  }
  if (p.out.pitch > 0 || p.out.ext[fa] % 2 == 0) return false;
  // a padded input: the same enumeration over the real elements as the
  // one-line kernel's; an unpadded one keeps its own
  q = p;
  if (p.pad_real > 0 && !lines_linear_unpadded(p, q)) return false;
  q.dual_split = 2;
  return true;
}

template <typename Real, int N, int RM>
hipError_t launch_dual(LinePass p, int B, hipStream_t s) {
  if constexpr (!dual_kernel_exists<Real, N>()) {
    return hipErrorInvalidConfiguration;
  } else {
    using Cfg = FastCfg<Real, N, RM>;
    constexpr int T = Cfg::T;
    const int P2 = B / 2;
    const size_t smem = static_cast<size_t>(B) * dual_line_bytes<Real, N, RM>();
    if (P2 * T > 1024 || smem > 160 * 1024) return hipErrorInvalidConfiguration;
    p.lines_per_block = B;
    static std::atomic<size_t> lds_limit[6] = {64 * 1024, 64 * 1024, 64 * 1024, 64 * 1024, 64 * 1024, 64 * 1024};
    const int nt = p.nontemporal ? 1 : 0;
    // 0: paired 16-byte loads, 1: adjacent input lines as two 8-byte loads,
    // 2: every line with its own addresses
    const int mode = p.dual_split & 2 ? 2 : (p.dual_split & 1);
    const void* fns[2][3] = {
        {reinterpret_cast<const void*>(&fft_line2_fast<Real, N, RM, false, 0>),
         reinterpret_cast<const void*>(&fft_line2_fast<Real, N, RM, false, 1>),
         reinterpret_cast<const void*>(&fft_line2_fast<Real, N, RM, false, 2>)},
        {reinterpret_cast<const void*>(&fft_line2_fast<Real, N, RM, true, 0>),
         reinterpret_cast<const void*>(&fft_line2_fast<Real, N, RM, true, 1>),
         reinterpret_cast<const void*>(&fft_line2_fast<Real, N, RM, true, 2>)}};
    const hipError_t e = raise_lds_limit(lds_limit[3 * nt + mode], fns[nt][mode], smem);
    if (e != hipSuccess) return e;
    const unsigned grid = static_cast<unsigned>((static_cast<int64_t>(p.nlines) + B - 1) / B);
    if (grid == 0) return hipSuccess;
    if (p.swizzle) {
      const int32_t inner = p.batch_ext[p.nbatch - 1];
      const int32_t C = (inner > 0 && inner % B == 0) ? inner / B : 0;
      p.swizzle = (C > 1 && grid % static_cast<unsigned>(C) == 0 && grid / C > 1) ? C : 0;
    }
    const dim3 g(grid), b(P2 * T);
    if (nt) {
      if (mode == 2) hipLaunchKernelGGL((fft_line2_fast<Real, N, RM, true, 2>), g, b, smem, s, p);
      else if (mode == 1) hipLaunchKernelGGL((fft_line2_fast<Real, N, RM, true, 1>), g, b, smem, s, p);
      else hipLaunchKernelGGL((fft_line2_fast<Real, N, RM, true, 0>), g, b, smem, s, p);
    } else {
      if (mode == 2) hipLaunchKernelGGL((fft_line2_fast<Real, N, RM, false, 2>), g, b, smem, s, p);
      else if (mode == 1) hipLaunchKernelGGL((fft_line2_fast<Real, N, RM, false, 1>), g, b, smem, s, p);
      else hipLaunchKernelGGL((fft_line2_fast<Real, N, RM, false, 0>), g, b, smem, s, p);
    }
    return hipGetLastError();
  }
}

template <typename Real, int N, int RM>
hipError_t launch_fast(LinePass p, hipStream_t s) {
  using Cfg = FastCfg<Real, N, RM>;
  using C = cplx_t<Real>;
  constexpr int T = Cfg::T;
  if constexpr (dual_kernel_exists<Real, N>()) {
    // default: 32 lines (256-byte row segments) or as many as 1024 lanes hold
    const int Bd = p.lines_per_block > 0 ? p.lines_per_block : (T <= 64 ? 32 : 2 * (1024 / T));
    const int mode = dual_lines_mode(p, Bd);
    if (mode >= 0) {
      p.dual_split = mode;
      return launch_dual<Real, N, RM>(p, Bd, s);
    }
    LinePass q;
    if (dual_lines_linear(p, Bd, q)) return launch_dual<Real, N, RM>(q, Bd, s);
  }
  {
    LinePass q;
    if (p.linear_lines && lines_linear_unpadded(p, q)) p = q;
  }
  const bool split = split_available<Real, N>() && p.split_lds != 0;
  const size_t line_bytes = split ? line_lds_bytes<Real, N, RM, true>() : line_lds_bytes<Real, N, RM, false>();
  int B = p.lines_per_block;
  if (B <= 0) {
    if (p.lines_fast) {
      B = 128 / static_cast<int>(sizeof(C));  // lines per 128-byte row segment
      while (B * T > 1024 && B > 1) B /= 2;
    } else {
      B = T >= 256 ? 1 : 256 / T;
    }
    while (B > 1 && static_cast<size_t>(B) * line_bytes > 160 * 1024) B /= 2;
  }
  const size_t smem = static_cast<size_t>(B) * line_bytes;
  if (B * T > 1024 || smem > 160 * 1024) return hipErrorInvalidConfiguration;
  p.lines_per_block = B;
  // two real lines per complex FFT for forward R2C and inverse C2R passes
  // (kernels for 8 values per thread)
  // (not for FP64 lengths with an odd factor: the pair kernel's two sets of
  // 20-40 complex values per lane spill registers there)
  const bool pair = FastVariant<Real, N, RM, false, false>::kPairs && RM == 8 && p.paired &&
                    ((p.kind == kR2C && !p.inverse) || (p.kind == kC2R && p.inverse));
  const bool nt = p.nontemporal != 0;
  const int v = 4 * (split ? 1 : 0) + 2 * (pair ? 1 : 0) + (nt ? 1 : 0);
  // raise the dynamic-LDS limit once per kernel instance
  static std::atomic<size_t> lds_limit[8] = {64 * 1024, 64 * 1024, 64 * 1024, 64 * 1024,
                                             64 * 1024, 64 * 1024, 64 * 1024, 64 * 1024};
  {
    const hipError_t e = raise_lds_limit(lds_limit[v], fast_variant_fn<Real, N, RM>(split, pair, nt), smem);
    if (e != hipSuccess) return e;
  }
  const int64_t units = pair ? (static_cast<int64_t>(p.nlines) + 1) / 2 : p.nlines;
  const unsigned grid = static_cast<unsigned>((units + B - 1) / B);
  if (grid == 0) return hipSuccess;
  if (p.swizzle) {
    // C = blocks per row of the fastest batch axis; the walk needs whole rows
    const int32_t inner = p.nbatch > 0 ? p.batch_ext[p.nbatch - 1] : 0;
    const int32_t per = pair ? 2 * B : B;
    const int32_t C = (inner > 0 && inner % per == 0) ? inner / per : 0;
    p.swizzle = (C > 1 && grid % static_cast<unsigned>(C) == 0 && grid / C > 1) ? C : 0;
  }
  if (nt) {
    launch_fast_variant<Real, N, RM, true>(split, pair, grid, static_cast<unsigned>(T * B), smem, s, p);
  } else {
    launch_fast_variant<Real, N, RM, false>(split, pair, grid, static_cast<unsigned>(T * B), smem, s, p);
  }
  return hipGetLastError();
}

template <typename Real, int N>
hipError_t launch_fast_r(const LinePass& p, hipStream_t s) {
  switch (p.values_per_thread) {
    case 4: return launch_fast<Real, N, 4>(p, s);
    case 16: return launch_fast<Real, N, 16>(p, s);
    default: return launch_fast<Real, N, 8>(p, s);
  }
}

// FP64 products pass: the paired R2C kernel (8 values per thread) with the
// products formed at the loads; block shape as launch_fast's pair path.
template <int N>
hipError_t launch_products(LinePass p, hipStream_t s) {
  using Real = double;
  using Cfg = FastCfg<Real, N, 8>;
  constexpr int T = Cfg::T;
  if (!p.products || p.kind != kR2C || p.inverse) return hipErrorInvalidValue;
  const bool split = p.split_lds != 0;
  const size_t line_bytes = split ? line_lds_bytes<Real, N, 8, true>() : line_lds_bytes<Real, N, 8, false>();
  // pairs per block: whole positions (3 pairs each) filling at least three
  // full waves (192 lanes; the 256^3 NS step took 8.01 ms with it, 8.23 ms
  // with the R2C pass's own 3 pairs of 32 lanes, 8.36-8.49 ms with 6 or 12
  // waves), cut to what this shape can hold
  int B = 3;
  while (B * T < 192 || (B * T) % 64) B += 3;
  while (B > 3 && (B * T > 1024 || static_cast<size_t>(B) * line_bytes > 160 * 1024)) B -= 3;
  const size_t smem = static_cast<size_t>(B) * line_bytes;
  if (B * T > 1024 || smem > 160 * 1024 || p.nlines % 6 != 0 || p.line_begin % 6 != 0)
    return hipErrorInvalidConfiguration;
  // the staged velocity lines (3 per position) fit in the three slots' LDS
  static_assert(3 * sizeof(Real) * N <= 3 * line_lds_bytes<Real, N, 8, true>(), "staging area");
  p.lines_per_block = B;
  p.swizzle = 0;
  const bool nt = p.nontemporal != 0;
  static std::atomic<size_t> lds_limit[4] = {64 * 1024, 64 * 1024, 64 * 1024, 64 * 1024};
  const void* fn = split ? (nt ? reinterpret_cast<const void*>(&fft_pair_fast<Real, N, 8, true, true, true>)
                               : reinterpret_cast<const void*>(&fft_pair_fast<Real, N, 8, false, true, true>))
                         : (nt ? reinterpret_cast<const void*>(&fft_pair_fast<Real, N, 8, true, false, true>)
                               : reinterpret_cast<const void*>(&fft_pair_fast<Real, N, 8, false, false, true>));
  const hipError_t e = raise_lds_limit(lds_limit[2 * (split ? 1 : 0) + (nt ? 1 : 0)], fn, smem);
  if (e != hipSuccess) return e;
  const int64_t units = (static_cast<int64_t>(p.nlines) + 1) / 2;
  const unsigned grid = static_cast<unsigned>((units + B - 1) / B);
  if (grid == 0) return hipSuccess;
  const dim3 g(grid), b(static_cast<unsigned>(T * B));
  if (split) {
    if (nt) hipLaunchKernelGGL((fft_pair_fast<Real, N, 8, true, true, true>), g, b, smem, s, p);
    else hipLaunchKernelGGL((fft_pair_fast<Real, N, 8, false, true, true>), g, b, smem, s, p);
  } else {
    if (nt) hipLaunchKernelGGL((fft_pair_fast<Real, N, 8, true, false, true>), g, b, smem, s, p);
    else hipLaunchKernelGGL((fft_pair_fast<Real, N, 8, false, false, true>), g, b, smem, s, p);
  }
  return hipGetLastError();
}

}  // namespace

// fast-kernel dispatch for one precision and a range of lengths (defined in
// fft_lines_*.hip); hipErrorNotSupported for lengths outside the range
hipError_t dispatch_fast_f64_short(const LinePass& p, hipStream_t s);
hipError_t dispatch_fast_f64_long(const LinePass& p, hipStream_t s);
hipError_t dispatch_fast_f32_short(const LinePass& p, hipStream_t s);
hipError_t dispatch_fast_f32_long(const LinePass& p, hipStream_t s);
hipError_t dispatch_fast_f64_xl(const LinePass& p, hipStream_t s);
hipError_t dispatch_fast_f32_xl(const LinePass& p, hipStream_t s);
hipError_t dispatch_products_f64(const LinePass& p, hipStream_t s);

}  // namespace dfft
