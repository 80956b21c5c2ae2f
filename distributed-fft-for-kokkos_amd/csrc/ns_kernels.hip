// Pointwise kernels of the pseudo-spectral Navier-Stokes solver.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "ns_kernels.hpp"

namespace dfft {

namespace {

__device__ __forceinline__ int signed_index(int g, int n) { return g < (n + 1) / 2 ? g : g - n; }

struct Mode {
  double kx, ky, kz, ksq, inv_ksq, mask;
  bool kz0;
};

// block-local (ix, iy, iz) of point m of a LayoutRight (ex, ey, ez) block
// (32-bit multiply-shift divisions; m < 2^31)
__device__ __forceinline__ void block_coords(const NsBlock& b, int64_t m, int& ix, int& iy, int& iz) {
  const uint32_t mm = static_cast<uint32_t>(m);
  const uint32_t r = b.div2.div(mm);
  iz = static_cast<int>(mm - r * static_cast<uint32_t>(b.ext[2]));
  const uint32_t q = b.div1.div(r);
  iy = static_cast<int>(r - q * static_cast<uint32_t>(b.ext[1]));
  ix = static_cast<int>(q);
}

// wavenumbers of spectral point m (LayoutRight (ex, ey, ez) block)
__device__ __forceinline__ Mode mode_of(const NsGrid& g, const NsBlock& b, int64_t m) {
  int ix, iy, iz;
  block_coords(b, m, ix, iy, iz);
  const int sx = signed_index(ix + b.start[0], g.n);
  const int sy = signed_index(iy + b.start[1], g.n);
  const int sz = iz + b.start[2];
  Mode md;
  md.kx = g.kscale * sx;
  md.ky = g.kscale * sy;
  md.kz = g.kscale * sz;
  md.ksq = md.kx * md.kx + md.ky * md.ky + md.kz * md.kz;
  md.inv_ksq = (sx == 0 && sy == 0 && sz == 0) ? 0.0 : 1.0 / md.ksq;
  const bool cut = (sx < 0 ? -sx : sx) > g.cutoff || (sy < 0 ? -sy : sy) > g.cutoff || sz > g.cutoff;
  md.mask = cut ? 0.0 : 1.0;
  md.kz0 = sz == 0;
  return md;
}

__device__ __forceinline__ double2 cscale(double2 a, double s) { return double2{a.x * s, a.y * s}; }
__device__ __forceinline__ double2 cadd2(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double2 ci(double2 a) { return double2{-a.y, a.x}; }  // i * a

__device__ __forceinline__ void project(const Mode& md, double2* v) {
  const double2 div = cscale(cadd2(cadd2(cscale(v[0], md.kx), cscale(v[1], md.ky)), cscale(v[2], md.kz)),
                             md.inv_ksq);
  v[0] = double2{v[0].x - md.kx * div.x, v[0].y - md.kx * div.y};
  v[1] = double2{v[1].x - md.ky * div.x, v[1].y - md.ky * div.y};
  v[2] = double2{v[2].x - md.kz * div.x, v[2].y - md.kz * div.y};
}

inline unsigned blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

}  // namespace

__global__ void __launch_bounds__(256) products_kernel(const double* __restrict__ u, double* __restrict__ p,
                                                       int64_t npts) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < npts; i += stride) {
    const double a = u[i], b = u[npts + i], c = u[2 * npts + i];
    p[i] = a * a;
    p[npts + i] = a * b;
    p[2 * npts + i] = a * c;
    p[3 * npts + i] = b * b;
    p[4 * npts + i] = b * c;
    p[5 * npts + i] = c * c;
  }
}

hipError_t ns_products(const double* u3, double* p6, int64_t npts, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  hipLaunchKernelGGL(products_kernel, dim3(blocks_for(npts)), dim3(256), 0, s, u3, p6, npts);
  return hipGetLastError();
}

// advection (main.cpp:716-738): f = [u, v, w, dudx, dudy, dudz, dvdx, ..., dwdz]
__global__ void __launch_bounds__(256) advection_kernel(const double* __restrict__ f, double* __restrict__ adv,
                                                        int64_t npts) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < npts; i += stride) {
    const double u = f[i], v = f[npts + i], w = f[2 * npts + i];
    for (int c = 0; c < 3; ++c) {
      const double* g = f + (3 + 3 * c) * npts;
      adv[c * npts + i] = u * g[i] + v * g[npts + i] + w * g[2 * npts + i];
    }
  }
}

hipError_t ns_advection(const double* f12, double* adv3, int64_t npts, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  hipLaunchKernelGGL(advection_kernel, dim3(blocks_for(npts)), dim3(256), 0, s, f12, adv3, npts);
  return hipGetLastError();
}

// dealiased state and its 9 spectral derivatives (main.cpp:576-597, 675-694)
__global__ void __launch_bounds__(256) gradients_kernel(const NsGrid g, const NsBlock b, const double2* y,
                                                        double2* f) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    for (int c = 0; c < 3; ++c) {
      const double2 v = cscale(y[c * m_tot + m], md.mask);
      f[c * m_tot + m] = v;
      const double2 iv = ci(v);
      f[(3 + 3 * c) * m_tot + m] = cscale(iv, md.kx);
      f[(4 + 3 * c) * m_tot + m] = cscale(iv, md.ky);
      f[(5 + 3 * c) * m_tot + m] = cscale(iv, md.kz);
    }
  }
}

hipError_t ns_gradients(const NsGrid& g, const NsBlock& b, const double2* y3, double2* f12, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(gradients_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, y3, f12);
  return hipGetLastError();
}

// Streaming (nontemporal) element access of the spectral fields: the stage
// update reads and writes 18-24 fields once each (2.4-3.2 GB at 256^3), far
// beyond the caches, so its lines need not displace anything (256^3 step
// 9.17 -> 8.89 ms; the 128^3 adjoint unchanged).
typedef double StreamD2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 stream_load(const double2* p) {
  const StreamD2 v = __builtin_nontemporal_load(reinterpret_cast<const StreamD2*>(p));
  return double2{v.x, v.y};
}
__device__ __forceinline__ void stream_store(double2* p, double2 x) {
  StreamD2 v;
  v.x = x.x;
  v.y = x.y;
  __builtin_nontemporal_store(v, reinterpret_cast<StreamD2*>(p));
}

__global__ void __launch_bounds__(256) rhs_rk_kernel(const NsGrid g, const NsBlock b, const NsStage st) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    double2 nl[3];
    if (st.form == 0) {
      // N_i = i sum_j k_j P_ij with P = [uu uv uw; uv vv vw; uw vw ww]
      const double2 p0 = stream_load(st.nl + m), p1 = stream_load(st.nl + (m_tot + m));
      const double2 p2 = stream_load(st.nl + (2 * m_tot + m)), p3 = stream_load(st.nl + (3 * m_tot + m));
      const double2 p4 = stream_load(st.nl + (4 * m_tot + m)), p5 = stream_load(st.nl + (5 * m_tot + m));
      nl[0] = ci(cadd2(cadd2(cscale(p0, md.kx), cscale(p1, md.ky)), cscale(p2, md.kz)));
      nl[1] = ci(cadd2(cadd2(cscale(p1, md.kx), cscale(p3, md.ky)), cscale(p4, md.kz)));
      nl[2] = ci(cadd2(cadd2(cscale(p2, md.kx), cscale(p4, md.ky)), cscale(p5, md.kz)));
    } else {
      nl[0] = stream_load(st.nl + m);
      nl[1] = stream_load(st.nl + (m_tot + m));
      nl[2] = stream_load(st.nl + (2 * m_tot + m));
    }
    // dealias, combine (-adv - nu k^2 u), project (main.cpp:655-665, 780-810)
    double2 y[3], r[3];
    const double visc = -g.nu * md.ksq;
    for (int c = 0; c < 3; ++c) {
      y[c] = cscale(stream_load(st.y + (c * m_tot + m)), md.mask);
      r[c] = double2{-md.mask * nl[c].x + visc * y[c].x, -md.mask * nl[c].y + visc * y[c].y};
    }
    project(md, r);
    // RK4th::advance (math_utils.hpp:96-137)
    for (int c = 0; c < 3; ++c) {
      const int64_t e = c * m_tot + m;
      const double2 kk = cscale(r[c], g.dt);
      double2 yn;
      if (st.stage == 0) {
        stream_store(st.yc + e, y[c]);
        stream_store(st.k1 + e, kk);
        yn = cadd2(y[c], cscale(kk, 0.5));
      } else if (st.stage == 1) {
        stream_store(st.k2 + e, kk);
        yn = cadd2(stream_load(st.yc + e), cscale(kk, 0.5));
      } else if (st.stage == 2) {
        stream_store(st.k3 + e, kk);
        yn = cadd2(stream_load(st.yc + e), kk);
      } else {
        const double2 k1 = stream_load(st.k1 + e), k2 = stream_load(st.k2 + e), k3 = stream_load(st.k3 + e);
        const double2 sum = double2{(k1.x + 2.0 * k2.x + 2.0 * k3.x + kk.x) / 6.0,
                                    (k1.y + 2.0 * k2.y + 2.0 * k3.y + kk.y) / 6.0};
        yn = cadd2(stream_load(st.yc + e), sum);
      }
      stream_store(st.y + e, cscale(yn, md.mask));
    }
  }
}

hipError_t ns_rhs_rk(const NsGrid& g, const NsBlock& b, const NsStage& st, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(rhs_rk_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, st);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) dealias_project_kernel(const NsGrid g, const NsBlock b, double2* y) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    double2 v[3];
    for (int c = 0; c < 3; ++c) v[c] = cscale(y[c * m_tot + m], md.mask);
    project(md, v);
    for (int c = 0; c < 3; ++c) y[c * m_tot + m] = v[c];
  }
}

hipError_t ns_dealias_project(const NsGrid& g, const NsBlock& b, double2* y3, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(dealias_project_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, y3);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) taylor_green_kernel(const NsBlock b, double h, double v0, double* u) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    int ix, iy, iz;
    block_coords(b, m, ix, iy, iz);
    const double x = h * (ix + b.start[0]), y = h * (iy + b.start[1]), z = h * (iz + b.start[2]);
    double sx, cx, sy, cy, sz, cz;
    sincos(x, &sx, &cx);
    sincos(y, &sy, &cy);
    sincos(z, &sz, &cz);
    u[m] = v0 * sx * cy * cz;
    u[m_tot + m] = -v0 * cx * sy * cz;
    u[2 * m_tot + m] = 0.0;
  }
}

hipError_t ns_taylor_green(const NsBlock& b, double h, double v0, double* u3, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(taylor_green_kernel, dim3(blocks_for(m)), dim3(256), 0, s, b, h, v0, u3);
  return hipGetLastError();
}

// block-level tree reduction in LDS; deterministic for a fixed launch shape
__global__ void __launch_bounds__(256) energy_kernel(const NsGrid g, const NsBlock b, const double2* y,
                                                     double* partials) {
  __shared__ double red[256];
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  double acc = 0.0;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const uint32_t mm = static_cast<uint32_t>(m);
    const int iz = static_cast<int>(mm - b.div2.div(mm) * static_cast<uint32_t>(b.ext[2]));
    const double f = (iz + b.start[2]) == 0 ? 1.0 : 2.0;
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      const double2 v = y[c * m_tot + m];
      s += v.x * v.x + v.y * v.y;
    }
    acc += 0.5 * f * s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0] * g.coef;
}

hipError_t ns_energy_partials(const NsGrid& g, const NsBlock& b, const double2* y3, double* partials,
                              int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(energy_kernel, dim3(nblocks), dim3(256), 0, s, g, b, y3, partials);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) sum_kernel(const double* partials, int n, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

hipError_t ns_sum(const double* partials, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, partials, n, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tangent and adjoint pieces of the divergence-form right-hand side.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) products_tangent_kernel(const double* __restrict__ u,
                                                               const double* __restrict__ du,
                                                               double* __restrict__ p, int64_t npts) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < npts; i += stride) {
    const double a = u[i], b = u[npts + i], c = u[2 * npts + i];
    const double da = du[i], db = du[npts + i], dc = du[2 * npts + i];
    p[i] = 2.0 * a * da;
    p[npts + i] = da * b + a * db;
    p[2 * npts + i] = da * c + a * dc;
    p[3 * npts + i] = 2.0 * b * db;
    p[4 * npts + i] = db * c + b * dc;
    p[5 * npts + i] = 2.0 * c * dc;
  }
}

hipError_t ns_products_tangent(const double* u3, const double* du3, double* p6, int64_t npts, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  hipLaunchKernelGGL(products_tangent_kernel, dim3(blocks_for(npts)), dim3(256), 0, s, u3, du3, p6, npts);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) rhs_only_kernel(const NsGrid g, const NsBlock b, const double2* nl,
                                                       const double2* yin, double2* out) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    const double2 p0 = nl[m], p1 = nl[m_tot + m], p2 = nl[2 * m_tot + m];
    const double2 p3 = nl[3 * m_tot + m], p4 = nl[4 * m_tot + m], p5 = nl[5 * m_tot + m];
    double2 n3[3];
    n3[0] = ci(cadd2(cadd2(cscale(p0, md.kx), cscale(p1, md.ky)), cscale(p2, md.kz)));
    n3[1] = ci(cadd2(cadd2(cscale(p1, md.kx), cscale(p3, md.ky)), cscale(p4, md.kz)));
    n3[2] = ci(cadd2(cadd2(cscale(p2, md.kx), cscale(p4, md.ky)), cscale(p5, md.kz)));
    const double visc = -g.nu * md.ksq;
    double2 r[3];
    for (int c = 0; c < 3; ++c) {
      const double2 y = cscale(yin[c * m_tot + m], md.mask);
      r[c] = double2{-md.mask * n3[c].x + visc * y.x, -md.mask * n3[c].y + visc * y.y};
    }
    project(md, r);
    for (int c = 0; c < 3; ++c) out[c * m_tot + m] = r[c];
  }
}

hipError_t ns_rhs_only(const NsGrid& g, const NsBlock& b, const double2* nl6, const double2* y3, double2* out3,
                       hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(rhs_only_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, nl6, y3, out3);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) adjoint_prepare_kernel(const NsGrid g, const NsBlock b, const double2* lam,
                                                              double2* mu, double2* sig) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    double2 v[3];
    for (int c = 0; c < 3; ++c) v[c] = lam[c * m_tot + m];
    project(md, v);
    for (int c = 0; c < 3; ++c) {
      v[c] = cscale(v[c], md.mask);
      mu[c * m_tot + m] = v[c];
    }
    const double k[3] = {md.kx, md.ky, md.kz};
    int f = 0;
    for (int i = 0; i < 3; ++i) {
      for (int j = i; j < 3; ++j) {
        // -i (k_j mu_i + k_i mu_j)
        const double2 s2 = cadd2(cscale(v[i], k[j]), cscale(v[j], k[i]));
        sig[f * m_tot + m] = double2{s2.y, -s2.x};
        ++f;
      }
    }
  }
}

hipError_t ns_adjoint_prepare(const NsGrid& g, const NsBlock& b, const double2* lam3, double2* mu3,
                              double2* sig6, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_prepare_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, lam3, mu3, sig6);
  return hipGetLastError();
}

// S is stored in the order (00, 01, 02, 11, 12, 22)
__global__ void __launch_bounds__(256) adjoint_contract_kernel(const double* __restrict__ u,
                                                               const double* __restrict__ S,
                                                               double* __restrict__ c, int64_t npts) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < npts; i += stride) {
    const double a = u[i], b = u[npts + i], w = u[2 * npts + i];
    const double s00 = S[i], s01 = S[npts + i], s02 = S[2 * npts + i];
    const double s11 = S[3 * npts + i], s12 = S[4 * npts + i], s22 = S[5 * npts + i];
    c[i] = a * s00 + b * s01 + w * s02;
    c[npts + i] = a * s01 + b * s11 + w * s12;
    c[2 * npts + i] = a * s02 + b * s12 + w * s22;
  }
}

hipError_t ns_adjoint_contract(const double* u3, const double* S6, double* c3, int64_t npts, hipStream_t s) {
  if (npts <= 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_contract_kernel, dim3(blocks_for(npts)), dim3(256), 0, s, u3, S6, c3, npts);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) adjoint_finish_kernel(const NsGrid g, const NsBlock b, const double2* chat,
                                                             const double2* mu, double2* out) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    const double visc = -g.nu * md.ksq;
    for (int c = 0; c < 3; ++c) {
      const double2 ch = chat[c * m_tot + m], u = mu[c * m_tot + m];
      out[c * m_tot + m] = double2{-md.mask * ch.x + visc * u.x, -md.mask * ch.y + visc * u.y};
    }
  }
}

hipError_t ns_adjoint_finish(const NsGrid& g, const NsBlock& b, const double2* chat3, const double2* mu3,
                             double2* out3, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_finish_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, chat3, mu3, out3);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) axpby_kernel(int64_t n, double a, double2* x, double b, const double2* y) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 xv = x[i], yv = y[i];
    x[i] = double2{a * xv.x + b * yv.x, a * xv.y + b * yv.y};
  }
}

hipError_t ns_axpby(int64_t n, double a, double2* x, double b, const double2* y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpby_kernel, dim3(blocks_for(n)), dim3(256), 0, s, n, a, x, b, y);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) dealias_kernel(const NsGrid g, const NsBlock b, const double2* src,
                                                      double2* dst) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    for (int c = 0; c < 3; ++c) dst[c * m_tot + m] = cscale(src[c * m_tot + m], md.mask);
  }
}

hipError_t ns_dealias(const NsGrid& g, const NsBlock& b, const double2* src3, double2* dst3, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(dealias_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, src3, dst3);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) scale_kernel(int64_t n, double a, double* x) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) x[i] *= a;
}

hipError_t ns_scale(int64_t n, double a, double* x, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(n)), dim3(256), 0, s, n, a, x);
  return hipGetLastError();
}

}  // namespace dfft

namespace dfft {

// ---------------------------------------------------------------------------
// Diagnostics (main.cpp:781-910): vorticity and the radial energy spectrum.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) vorticity_kernel(const NsGrid g, const NsBlock b, const double2* y,
                                                        double2* w) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    const double2 u = y[m], v = y[m_tot + m], z = y[2 * m_tot + m];
    // i (ky w - kz v), i (kz u - kx w), i (kx v - ky u)
    w[m] = ci(double2{md.ky * z.x - md.kz * v.x, md.ky * z.y - md.kz * v.y});
    w[m_tot + m] = ci(double2{md.kz * u.x - md.kx * z.x, md.kz * u.y - md.kx * z.y});
    w[2 * m_tot + m] = ci(double2{md.kx * v.x - md.ky * u.x, md.kx * v.y - md.ky * u.y});
  }
}

hipError_t ns_vorticity(const NsGrid& g, const NsBlock& b, const double2* y3, double2* w3, hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(vorticity_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, y3, w3);
  return hipGetLastError();
}

// Each block bins its points into LDS (bins [k_i, k_{i+1}), the reference's
// half-open test) and writes one row of per-bin partial sums.
__global__ void __launch_bounds__(256) spectrum_kernel(const NsGrid g, const NsBlock b, const double2* y,
                                                       const double* bins, int nb, double* partials) {
  __shared__ double acc[kMaxSpectrumBins];
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const double width = bins[nb] / nb;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    const double k = sqrt(md.ksq);
    int i = width > 0.0 ? static_cast<int>(k / width) : 0;
    i = i < 0 ? 0 : (i > nb - 1 ? nb - 1 : i);
    while (i > 0 && k < bins[i]) --i;
    while (i < nb - 1 && k >= bins[i + 1]) ++i;
    if (!(k >= bins[i] && k < bins[i + 1])) continue;
    double e = 0.0;
    for (int c = 0; c < 3; ++c) {
      const double2 v = y[c * m_tot + m];
      e += v.x * v.x + v.y * v.y;
    }
    atomicAdd(&acc[i], 0.5 * (md.kz0 ? 1.0 : 2.0) * e);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += blockDim.x) partials[static_cast<int64_t>(blockIdx.x) * nb + i] = acc[i];
}

hipError_t ns_energy_spectrum_partials(const NsGrid& g, const NsBlock& b, const double2* y3, const double* bins,
                                       int nb, double* partials, int nblocks, hipStream_t s) {
  if (nb < 1 || nb > kMaxSpectrumBins) return hipErrorInvalidValue;
  hipLaunchKernelGGL(spectrum_kernel, dim3(nblocks), dim3(256), 0, s, g, b, y3, bins, nb, partials);
  return hipGetLastError();
}

// out[i] = sum over rows r of partials[r * nb + i], rows in order
__global__ void __launch_bounds__(256) bin_sum_kernel(const double* partials, int rows, int nb, double* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int r = 0; r < rows; ++r) a += partials[static_cast<int64_t>(r) * nb + i];
    out[i] = a;
  }
}

hipError_t ns_bin_sum(const double* partials, int rows, int nb, double* out, hipStream_t s) {
  hipLaunchKernelGGL(bin_sum_kernel, dim3(1), dim3(256), 0, s, partials, rows, nb, out);
  return hipGetLastError();
}

}  // namespace dfft

namespace dfft {

// ---------------------------------------------------------------------------
// Reverse RK4 step: the adjoint's finishing step fused with the stage
// combination.  With t = -D chat - nu k^2 mu (= R'(y)^* of the stage seed):
//   mode 1: acc += t;  fb = dt (a lam + c t)   (the next stage's seed)
//   mode 2: lam = D (acc + t)                  (the adjoint of the step input)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) adjoint_finish_rk_kernel(const NsGrid g, const NsBlock b, const double2* chat,
                                                                const double2* mu, double2* acc, double2* lam,
                                                                double2* fb, double a, double c, int mode) {
  const int64_t m_tot = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < m_tot; m += stride) {
    const Mode md = mode_of(g, b, m);
    const double visc = -g.nu * md.ksq;
    for (int k = 0; k < 3; ++k) {
      const int64_t e = k * m_tot + m;
      const double2 ch = chat[e], u = mu[e];
      const double2 t{-md.mask * ch.x + visc * u.x, -md.mask * ch.y + visc * u.y};
      const double2 s = acc[e];
      const double2 s2{s.x + t.x, s.y + t.y};
      if (mode == 1) {
        acc[e] = s2;
        const double2 l = lam[e];
        fb[e] = double2{g.dt * (a * l.x + c * t.x), g.dt * (a * l.y + c * t.y)};
      } else {
        lam[e] = cscale(s2, md.mask);
      }
    }
  }
}

hipError_t ns_adjoint_finish_rk(const NsGrid& g, const NsBlock& b, const double2* chat3, const double2* mu3,
                                double2* acc3, double2* lam3, double2* fb3, double a, double c, int mode,
                                hipStream_t s) {
  const int64_t m = static_cast<int64_t>(b.ext[0]) * b.ext[1] * b.ext[2];
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_finish_rk_kernel, dim3(blocks_for(m)), dim3(256), 0, s, g, b, chat3, mu3, acc3, lam3,
                     fb3, a, c, mode);
  return hipGetLastError();
}

// fb = c lam, acc = lam (start of a reverse RK4 step)
__global__ void __launch_bounds__(256) rk_adjoint_init_kernel(int64_t n, const double2* lam, double2* fb,
                                                              double2* acc, double c) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 l = lam[i];
    fb[i] = double2{c * l.x, c * l.y};
    acc[i] = l;
  }
}

hipError_t ns_rk_adjoint_init(int64_t n, const double2* lam, double2* fb, double2* acc, double c, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(rk_adjoint_init_kernel, dim3(blocks_for(n)), dim3(256), 0, s, n, lam, fb, acc, c);
  return hipGetLastError();
}

}  // namespace dfft
