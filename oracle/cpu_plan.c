/* Oracle (test infrastructure only): OpenMP CPU executor of the reference's
 * slab/pencil block sequence, used as the timed CPU baseline beside the GPU
 * numbers and pinned against the Python oracle plan (oracle/plan_exec.py) on
 * the same inputs.
 *
 * It runs the same steps as the reference's CPU path (Kokkos OpenMP +
 * FFTW, KokkosFFT_Distributed_SlabPlan.hpp:1143-1267, PencilPlan.hpp:
 * 1444-1598), for P virtual ranks held in one process:
 *   - FFT blocks: batched 1-D C2C transforms along one axis of each rank's
 *     local array (cpu_fft_axis), out of place on the first axis pass;
 *   - Transpose blocks: pack (PackUnpack.hpp:217-233) -> all-to-all of the
 *     (P, e...) buffers (MPI_All2All.hpp:54-57) -> unpack (:385-417);
 *   - local transposes (kokkos-fft Impl::transpose) and the separate
 *     normalisation sweep (SlabPlan.hpp:1344-1358).
 * Every entry point takes one pointer per virtual rank and works on all
 * ranks in one OpenMP parallel region, so P ranks of T/P threads each (the
 * reference's mpirun -np P with OMP_NUM_THREADS = cores/P) become one team
 * of T threads; a call returns when every rank is done (the fence and the
 * blocking MPI_Alltoall of TransBlock.hpp:97-106).
 *
 * FFTs of power-of-two length use an iterative radix-4/2 Stockham
 * decimation-in-frequency transform (natural-order output) whose twiddles
 * are evaluated in long double and rounded once; other lengths use the
 * recursive transform of fft_ref.c.  Strided axes are processed 8 lines at
 * a time: the chunk is gathered as 8 interleaved sequences (128-byte rows),
 * transformed together (the Stockham stride starts at 8) and scattered back.
 *
 * Arrays are complex double, LayoutRight (C order), D <= 4.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cx;

void fft_c2c(long n, long nlines, const double* in, double* out, int sign); /* fft_ref.c */

static const long double TWO_PI = 6.283185307179586476925286766559005768L;
#define VLINES 8

void cpu_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int cpu_get_threads(void) { return omp_get_max_threads(); }

static int is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

static long prod(const long* e, int lo, int hi) {
  long p = 1;
  for (int i = lo; i < hi; ++i) p *= e[i];
  return p;
}

/* tw[e] = exp(sign 2 pi i e / n), e in [0, n) */
static cx* make_twiddles(long n, int sign) {
  cx* tw = (cx*)malloc(sizeof(cx) * (size_t)n);
  for (long e = 0; e < n; ++e) {
    long double a = TWO_PI * (long double)e / (long double)n;
    tw[e].re = (double)cosl(a);
    tw[e].im = (double)((long double)sign * sinl(a));
  }
  return tw;
}

static inline cx cmul(cx a, cx b) {
  cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}

/* V interleaved transforms of length n (power of two): element k of sequence
 * c at x[k * V + c].  y is scratch of the same size.  Returns the buffer that
 * holds the result (x or y). */
static cx* stockham(long n, long V, cx* x, cx* y, const cx* tw, int sign) {
  long L = n, s = V;
  while (L > 1) {
    const long step = n / L;
    if (L % 4 == 0) {
      const long m = L / 4;
      for (long p = 0; p < m; ++p) {
        const cx w1 = tw[p * step], w2 = tw[2 * p * step], w3 = tw[3 * p * step];
        const cx* xa = x + s * p;
        const cx* xb = x + s * (p + m);
        const cx* xc = x + s * (p + 2 * m);
        const cx* xd = x + s * (p + 3 * m);
        cx* y0 = y + s * (4 * p);
        cx* y1 = y0 + s;
        cx* y2 = y1 + s;
        cx* y3 = y2 + s;
        for (long q = 0; q < s; ++q) {
          const cx a = xa[q], b = xb[q], c = xc[q], d = xd[q];
          const cx apc = {a.re + c.re, a.im + c.im}, amc = {a.re - c.re, a.im - c.im};
          const cx bpd = {b.re + d.re, b.im + d.im}, bmd = {b.re - d.re, b.im - d.im};
          /* jb = sign * i * (b - d) */
          const cx jb = sign < 0 ? (cx){bmd.im, -bmd.re} : (cx){-bmd.im, bmd.re};
          y0[q] = (cx){apc.re + bpd.re, apc.im + bpd.im};
          y1[q] = cmul(w1, (cx){amc.re + jb.re, amc.im + jb.im});
          y2[q] = cmul(w2, (cx){apc.re - bpd.re, apc.im - bpd.im});
          y3[q] = cmul(w3, (cx){amc.re - jb.re, amc.im - jb.im});
        }
      }
      L = m;
      s *= 4;
    } else { /* L == 2 */
      for (long q = 0; q < s; ++q) {
        const cx a = x[q], b = x[q + s];
        y[q] = (cx){a.re + b.re, a.im + b.im};
        y[q + s] = (cx){a.re - b.re, a.im - b.im};
      }
      L = 1;
      s *= 2;
    }
    cx* t = x;
    x = y;
    y = t;
  }
  return x;
}

/* Unscaled C2C along `axis` of every rank's array (sign -1 forward, +1
 * backward); out may equal in.  ext holds P rows of D extents. */
int cpu_fft_axis(int P, void** ins, void** outs, const long* ext, int D, int axis, int sign) {
  long n = 0;
  for (int r = 0; r < P; ++r) {
    const long m = ext[r * D + axis];
    if (prod(ext + r * D, 0, D) == 0) continue;
    if (n != 0 && m != n) return -1; /* FFT axes are whole on every rank */
    n = m;
  }
  if (n == 0) return 0;
  const int fast = is_pow2(n);
  cx* tw = fast ? make_twiddles(n, sign) : NULL;
  /* work items: (rank, outer index, chunk of inner indices) */
  long* first = (long*)malloc(sizeof(long) * (size_t)(P + 1));
  first[0] = 0;
  for (int r = 0; r < P; ++r) {
    const long* e = ext + r * D;
    long outer = prod(e, 0, axis), inner = prod(e, axis + 1, D);
    long chunks = inner == 1 ? 1 : (inner + VLINES - 1) / VLINES;
    first[r + 1] = first[r] + (prod(e, 0, D) ? outer * chunks : 0);
  }
  const long total = first[P];
#pragma omp parallel
  {
    cx* b0 = (cx*)malloc(sizeof(cx) * (size_t)(n * VLINES));
    cx* b1 = (cx*)malloc(sizeof(cx) * (size_t)(n * VLINES));
#pragma omp for schedule(static)
    for (long w = 0; w < total; ++w) {
      int r = 0;
      while (w >= first[r + 1]) ++r;
      const long* e = ext + r * D;
      const long inner = prod(e, axis + 1, D);
      const long chunks = inner == 1 ? 1 : (inner + VLINES - 1) / VLINES;
      const long o = (w - first[r]) / chunks, ch = (w - first[r]) % chunks;
      const cx* src = (const cx*)ins[r] + o * n * inner;
      cx* dst = (cx*)outs[r] + o * n * inner;
      if (inner == 1) { /* contiguous line */
        memcpy(b0, src, sizeof(cx) * (size_t)n);
        cx* res;
        if (fast) {
          res = stockham(n, 1, b0, b1, tw, sign);
        } else {
          fft_c2c(n, 1, (const double*)b0, (double*)b1, sign);
          res = b1;
        }
        memcpy(dst, res, sizeof(cx) * (size_t)n);
      } else {
        const long i0 = ch * VLINES;
        const long v = inner - i0 < VLINES ? inner - i0 : VLINES;
        for (long k = 0; k < n; ++k) {
          const cx* row = src + k * inner + i0;
          for (long c = 0; c < v; ++c) b0[k * v + c] = row[c];
        }
        cx* res;
        if (fast) {
          res = stockham(n, v, b0, b1, tw, sign);
        } else { /* de-interleave for the recursive transform */
          for (long c = 0; c < v; ++c)
            for (long k = 0; k < n; ++k) b1[c * n + k] = b0[k * v + c];
          fft_c2c(n, v, (const double*)b1, (double*)b0, sign);
          for (long c = 0; c < v; ++c)
            for (long k = 0; k < n; ++k) b1[k * v + c] = b0[c * n + k];
          res = b1;
        }
        for (long k = 0; k < n; ++k) {
          cx* row = dst + k * inner + i0;
          for (long c = 0; c < v; ++c) row[c] = res[k * v + c];
        }
      }
    }
    free(b0);
    free(b1);
  }
  free(first);
  free(tw);
  return 0;
}

/* bin_mapping (PackUnpack.hpp:40-46) */
static void bin_of(long n, long nbins, long ibin, long* start, long* len) {
  long base = n / nbins, rem = n % nbins;
  *len = base + (ibin < rem ? 1 : 0);
  *start = ibin * base + (ibin < rem ? ibin : rem);
}

/* Strides (in elements) of logical axis i of an array stored mapped by m
 * (storage position k holds logical axis m[k]) with storage extents e. */
static void logical_strides(const long* e, const int* m, int D, long* lstride, long* lext) {
  long s = 1;
  for (int k = D - 1; k >= 0; --k) {
    lstride[m[k]] = s;
    lext[m[k]] = e[k];
    s *= e[k];
  }
}

/* pack (unpack == 0) or unpack (unpack == 1) between every rank's mapped
 * array (storage extents arr_ext, P rows of D) and its (P, e_0..e_{D-1})
 * buffer (extents buf_ext, D + 1 values, logical order). */
static void pack_unpack(int P, void** arrs, const long* arr_ext, int D, void** bufs, const long* buf_ext,
                        const int* maps, int axis, int unpack) {
  const long np = buf_ext[0];
  const long* be = buf_ext + 1;
  const long rows = prod(be, 0, D - 1); /* buffer rows of the last logical axis */
  const long total = (long)P * np * rows;
#pragma omp parallel for schedule(static)
  for (long w = 0; w < total; ++w) {
    const int r = (int)(w / (np * rows));
    const long p = (w / rows) % np, row = w % rows;
    long lstride[4], lext[4];
    logical_strides(arr_ext + r * D, maps + r * D, D, lstride, lext);
    long start, len;
    bin_of(lext[axis], np, p, &start, &len);
    /* logical indices of the row's first element */
    long j[4], rem = row;
    for (int i = D - 2; i >= 0; --i) {
      j[i] = rem % be[i];
      rem /= be[i];
    }
    cx* buf = (cx*)bufs[r] + (p * rows + row) * be[D - 1];
    cx* arr = (cx*)arrs[r];
    long off = 0;
    int valid = 1;
    for (int i = 0; i < D - 1; ++i) {
      long l = j[i];
      if (i == axis) {
        if (l >= len) valid = 0;
        l += start;
      }
      if (l >= lext[i]) valid = 0;
      off += l * lstride[i];
    }
    const int last = D - 1;
    long lo = 0, hi = be[last], shift = 0;
    if (last == axis) {
      hi = hi < len ? hi : len;
      shift = start;
    }
    if (hi > lext[last] - shift) hi = lext[last] - shift;
    if (hi < 0) hi = 0;
    if (!valid) hi = 0;
    const long st = lstride[last];
    if (!unpack) {
      for (long k = lo; k < hi; ++k) buf[k] = arr[off + (k + shift) * st];
      for (long k = hi; k < be[last]; ++k) buf[k] = (cx){0.0, 0.0};
    } else {
      for (long k = lo; k < hi; ++k) arr[off + (k + shift) * st] = buf[k];
    }
  }
}

void cpu_pack(int P, void** srcs, const long* src_ext, int D, void** bufs, const long* buf_ext, const int* src_maps,
              int axis) {
  pack_unpack(P, srcs, src_ext, D, bufs, buf_ext, src_maps, axis, 0);
}

void cpu_unpack(int P, void** dsts, const long* dst_ext, int D, void** bufs, const long* buf_ext,
                const int* dst_maps, int axis) {
  pack_unpack(P, dsts, dst_ext, D, bufs, buf_ext, dst_maps, axis, 1);
}

/* MPI_Alltoall inside groups of `gsize` ranks (groups[g * gsize + i] = rank
 * of member i): member i's block p lands at member p as block i.  block =
 * elements per block. */
void cpu_all2all(int ngroups, int gsize, const int* groups, void** sends, void** recvs, long block) {
  const long total = (long)ngroups * gsize * gsize;
#pragma omp parallel for schedule(static)
  for (long w = 0; w < total; ++w) {
    const long g = w / ((long)gsize * gsize), i = (w / gsize) % gsize, p = w % gsize;
    const int src = groups[g * gsize + i], dst = groups[g * gsize + p];
    memcpy((cx*)recvs[dst] + i * block, (const cx*)sends[src] + p * block, sizeof(cx) * (size_t)block);
  }
}

/* out = transpose(in, map): out.extent(k) = in.extent(map[k]), out(j) = in(i)
 * with i[map[k]] = j_k (kokkos-fft Impl::transpose); in_ext P rows of D. */
void cpu_transpose(int P, void** ins, void** outs, const long* in_ext, int D, const int* map) {
  long first[65];
  if (P > 64) return;
  first[0] = 0;
  for (int r = 0; r < P; ++r) {
    const long* e = in_ext + r * D;
    long oe_last = e[map[D - 1]];
    long rows = oe_last ? prod(e, 0, D) / oe_last : 0;
    first[r + 1] = first[r] + rows;
  }
  const long total = first[P];
#pragma omp parallel for schedule(static)
  for (long w = 0; w < total; ++w) {
    int r = 0;
    while (w >= first[r + 1]) ++r;
    const long* e = in_ext + r * D;
    long istride[4], s = 1;
    for (int k = D - 1; k >= 0; --k) {
      istride[k] = s;
      s *= e[k];
    }
    long oe[4];
    for (int k = 0; k < D; ++k) oe[k] = e[map[k]];
    long row = w - first[r], rem = row, off = 0;
    for (int k = D - 2; k >= 0; --k) {
      long jk = rem % oe[k];
      rem /= oe[k];
      off += jk * istride[map[k]];
    }
    const cx* src = (const cx*)ins[r];
    cx* dst = (cx*)outs[r] + row * oe[D - 1];
    const long st = istride[map[D - 1]];
    for (long k = 0; k < oe[D - 1]; ++k) dst[k] = src[off + k * st];
  }
}

/* a *= s over n[r] complex elements of every rank */
void cpu_scale(int P, void** arrs, const long* n, double s) {
  for (int r = 0; r < P; ++r) {
    double* a = (double*)arrs[r];
    const long m = 2 * n[r];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < m; ++i) a[i] *= s;
  }
}

/* dst = src over n[r] complex elements of every rank */
void cpu_copy(int P, void** srcs, void** dsts, const long* n) {
  for (int r = 0; r < P; ++r) {
    const long m = n[r];
    const long chunk = 1 << 16;
    const long nch = (m + chunk - 1) / chunk;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nch; ++c) {
      long lo = c * chunk, len = m - lo < chunk ? m - lo : chunk;
      memcpy((cx*)dsts[r] + lo, (const cx*)srcs[r] + lo, sizeof(cx) * (size_t)len);
    }
  }
}

/* first touch of every page by the thread team (page placement as the
 * OpenMP loops above will use it) */
void cpu_zero(int P, void** arrs, const long* n) {
  for (int r = 0; r < P; ++r) {
    double* a = (double*)arrs[r];
    const long m = 2 * n[r];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < m; ++i) a[i] = 0.0;
  }
}
