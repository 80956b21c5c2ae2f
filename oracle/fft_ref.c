/* Oracle (test infrastructure only): CPU restatement of the discrete Fourier
 * transforms that the reference delegates to kokkos-fft (FFTW on the host,
 * rocFFT/hipFFT on AMD GPUs).  The reference calls
 * KokkosFFT::execute(DynPlan, in, out, Normalization::none) at
 * KokkosFFT_Distributed_SlabPlan.hpp:978-1081 and PencilPlan.hpp:1215-1348:
 *   forward  X[k] = sum_j x[j] exp(-2 pi i j k / n)
 *   backward x[j] = sum_k X[k] exp(+2 pi i j k / n)      (unscaled)
 * R2C keeps k = 0..n/2; C2R reads k = 0..n/2 and treats the spectrum as
 * Hermitian (imaginary parts of the k = 0 and k = n/2 terms are ignored).
 *
 * Two independent evaluations are provided:
 *   dft_ld_*  : direct O(n^2) sums in long double, phase reduced exactly as
 *               (j*k mod n) before the trig call (used to pin the fast path);
 *   fft_*     : recursive mixed-radix (smallest prime factor first) Cooley-
 *               Tukey in double with twiddles evaluated in long double and
 *               rounded once.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this
 * library, and only as the checker.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

static const long double TWO_PI_L = 6.283185307179586476925286766559005768L;

/* exp(sign * 2 pi i * e / n) with e already reduced to [0, n). */
static void twiddle_ld(long n, long e, int sign, long double* c, long double* s) {
  long double a = TWO_PI_L * (long double)e / (long double)n;
  *c = cosl(a);
  *s = (long double)sign * sinl(a);
}

/* ---------------- direct DFT in long double ---------------- */
void dft_ld_c2c(long n, long nlines, const double* in, double* out, int sign) {
  long double* cs = (long double*)malloc(sizeof(long double) * 2 * n);
  for (long e = 0; e < n; ++e) twiddle_ld(n, e, sign, &cs[2 * e], &cs[2 * e + 1]);
  for (long l = 0; l < nlines; ++l) {
    const double* x = in + 2 * n * l;
    double* y = out + 2 * n * l;
    for (long k = 0; k < n; ++k) {
      long double re = 0.0L, im = 0.0L;
      for (long j = 0; j < n; ++j) {
        long e = (long)(((long long)j * k) % n);
        long double c = cs[2 * e], s = cs[2 * e + 1];
        re += x[2 * j] * c - x[2 * j + 1] * s;
        im += x[2 * j] * s + x[2 * j + 1] * c;
      }
      y[2 * k] = (double)re;
      y[2 * k + 1] = (double)im;
    }
  }
  free(cs);
}

/* ---------------- recursive mixed-radix FFT in double ---------------- */
typedef struct {
  long n;        /* top-level length */
  cplx* tw;      /* tw[e] = exp(sign 2 pi i e / n), e in [0, n) */
  cplx* scratch; /* n elements */
} fft_ctx;

static long smallest_factor(long n) {
  if (n % 4 == 0 && n > 4) return 4;
  for (long p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

/* out[0..n) = DFT_n of in[0], in[s], ..., in[(n-1)s]; twiddle stride ts = N/n */
static void fft_rec(const fft_ctx* c, long n, const cplx* in, long s, cplx* out, long ts) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  long p = smallest_factor(n);
  long m = n / p;
  if (m == 1) { /* prime (or 4): direct */
    for (long k = 0; k < n; ++k) {
      double re = 0.0, im = 0.0;
      for (long j = 0; j < n; ++j) {
        long e = ((j * k) % n) * ts;
        cplx w = c->tw[e];
        re += in[j * s].re * w.re - in[j * s].im * w.im;
        im += in[j * s].re * w.im + in[j * s].im * w.re;
      }
      out[k].re = re;
      out[k].im = im;
    }
    return;
  }
  for (long r = 0; r < p; ++r) fft_rec(c, m, in + r * s, s * p, out + r * m, ts * p);
  cplx tmp[64];
  cplx* t = p <= 64 ? tmp : (cplx*)malloc(sizeof(cplx) * p);
  for (long k1 = 0; k1 < m; ++k1) {
    for (long r = 0; r < p; ++r) {
      /* twiddle W_n^{r k1} then a length-p DFT over r */
      cplx y = out[r * m + k1];
      long e = ((r * k1) % n) * ts;
      cplx w = c->tw[e];
      t[r].re = y.re * w.re - y.im * w.im;
      t[r].im = y.re * w.im + y.im * w.re;
    }
    for (long q = 0; q < p; ++q) {
      double re = 0.0, im = 0.0;
      for (long r = 0; r < p; ++r) {
        long e = ((r * q) % p) * (ts * m);
        cplx w = c->tw[e];
        re += t[r].re * w.re - t[r].im * w.im;
        im += t[r].re * w.im + t[r].im * w.re;
      }
      c->scratch[k1 + q * m].re = re;
      c->scratch[k1 + q * m].im = im;
    }
  }
  memcpy(out, c->scratch, sizeof(cplx) * n);
  if (t != tmp) free(t);
}

static void ctx_init(fft_ctx* c, long n, int sign) {
  c->n = n;
  c->tw = (cplx*)malloc(sizeof(cplx) * n);
  c->scratch = (cplx*)malloc(sizeof(cplx) * n);
  for (long e = 0; e < n; ++e) {
    long double co, si;
    twiddle_ld(n, e, sign, &co, &si);
    c->tw[e].re = (double)co;
    c->tw[e].im = (double)si;
  }
}

static void ctx_free(fft_ctx* c) {
  free(c->tw);
  free(c->scratch);
}

/* scratch is reused at every recursion level; each level writes the whole
 * of its own sub-range of `out` after all children finished, so a shared
 * n-element scratch suffices when children are processed sequentially. */
void fft_c2c(long n, long nlines, const double* in, double* out, int sign) {
  fft_ctx c;
  ctx_init(&c, n, sign);
  cplx* line = (cplx*)malloc(sizeof(cplx) * n);
  for (long l = 0; l < nlines; ++l) {
    memcpy(line, in + 2 * n * l, sizeof(cplx) * n);
    fft_rec(&c, n, line, 1, (cplx*)(out + 2 * n * l), 1);
  }
  free(line);
  ctx_free(&c);
}

/* real n -> complex n/2+1 per line */
void fft_r2c(long n, long nlines, const double* in, double* out) {
  fft_ctx c;
  ctx_init(&c, n, -1);
  cplx* line = (cplx*)malloc(sizeof(cplx) * n);
  cplx* res = (cplx*)malloc(sizeof(cplx) * n);
  long nh = n / 2 + 1;
  for (long l = 0; l < nlines; ++l) {
    for (long j = 0; j < n; ++j) {
      line[j].re = in[n * l + j];
      line[j].im = 0.0;
    }
    fft_rec(&c, n, line, 1, res, 1);
    memcpy(out + 2 * nh * l, res, sizeof(cplx) * nh);
  }
  free(line);
  free(res);
  ctx_free(&c);
}

/* complex n/2+1 -> real n per line (Hermitian extension, unscaled) */
void fft_c2r(long n, long nlines, const double* in, double* out) {
  fft_ctx c;
  ctx_init(&c, n, +1);
  cplx* line = (cplx*)malloc(sizeof(cplx) * n);
  cplx* res = (cplx*)malloc(sizeof(cplx) * n);
  long nh = n / 2 + 1;
  for (long l = 0; l < nlines; ++l) {
    const cplx* X = (const cplx*)(in + 2 * nh * l);
    for (long k = 0; k < n; ++k) {
      if (k < nh) {
        line[k] = X[k];
      } else {
        line[k].re = X[n - k].re;
        line[k].im = -X[n - k].im;
      }
    }
    line[0].im = 0.0;
    if (n % 2 == 0) line[n / 2].im = 0.0;
    fft_rec(&c, n, line, 1, res, 1);
    for (long j = 0; j < n; ++j) out[n * l + j] = res[j].re;
  }
  free(line);
  free(res);
  ctx_free(&c);
}
